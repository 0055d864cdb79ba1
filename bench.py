"""Benchmark of the element-assembly hot path on one MI355X.

A step is one residual + CSR Jacobian assembly (the residual and stiffness calls of one Newton iteration,
src/Solvers.jl:133-141) over the whole mesh, inputs resident in HBM: DOF update, FP32 element math, FP64 sums into
the residual field and the CSR values. Default workload: compressible neo-Hookean on the unit cube, 96^3 hex8
(884,736 elements, 2,738,019 DOFs, 217,238,121 CSR values), lambda = 1, mu = 0.5,
u = 0.05 sum_k a_k sin(k pi x) sin(k pi y) sin(k pi z).

Prints one JSON line: elements/s, ms/step, the step's HBM roofline fraction priced on the compulsory bytes per element
(HIP-event kernel timing on the assembler's stream), relative errors of every output (residual, CSR values,
matrix-free and assembled actions) against the FP64 oracle at the full size, and the oracle's CPU throughput on a
bounded sample (cpu_baseline, kind "port": the C++ FP64 restatement of the reference's CPU algorithm).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

METRIC = "residual+Jacobian elements/sec: FP32/learned vs FP64 plain at stated rel err"
HBM_PEAK_GBS = 8000.0

# per-config full sizes and the bounded sample used for parity and the CPU baseline
SIZES = {"poisson_hex8": (128, 56), "neohookean_hex8": (96, 32), "j2_hex8": (64, 28), "neohookean_tet10": (80, 14),
         "poisson_quad4": (64, 64)}
# algorithmic (compulsory) bytes per element for residual + tangent (int32 conn, FP32 coords, FP64 u, FP64 outputs written
# once, int32 COO->CSR map of NDOF^2 per element, FP32 state)
STEP_BYTES = {"poisson_hex8": 534.4, "neohookean_hex8": 4362.2, "j2_hex8": 4821.4, "neohookean_tet10": 6502.3}
# the same model for the residual alone and for the matrix-free action (FP64 v read as well)
RESIDUAL_BYTES = {"poisson_hex8": 60.7, "neohookean_hex8": 93.9, "j2_hex8": 542.9, "neohookean_tet10": 121.5}
MF_ACTION_BYTES = {"poisson_hex8": 68.9, "neohookean_hex8": 118.7, "j2_hex8": 568.0, "neohookean_tet10": 154.1}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_assembler(fec, wl, device=0):
    V = fec.FunctionSpace(wl.mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=device)
    bcs = []
    if len(wl.dirichlet_dofs):
        nodes = np.unique((wl.dirichlet_dofs - 1) // wl.nf + 1)
        wl.mesh.nodeset_nodes["_bc"] = nodes
        bcs = [fec.DirichletBC(n, lambda X, t: 0.0, nset_name="_bc") for n in var.names]
    p = fec.create_parameters(wl.mesh, asm, wl.physics, dirichlet_bcs=bcs)
    return asm, p


def kernel_bytes(wl, asm, with_r=True, with_k=True):
    """Algorithmic HBM bytes per launch of each kernel of the two-phase assembly (compulsory traffic)."""
    ne, nnpe = wl.mesh.conn.shape
    nn, nf = wl.mesh.num_nodes, wl.nf
    nd = wl.mesh.coords.shape[1]
    ndof = nnpe * nf
    npairs = nnpe * (nnpe + 1) // 2
    vstride = (nnpe + 3) // 4 * 4
    state = 2 * 4 * 7 * 8 * ne if wl.physics.kind == 3 else 0
    kel = 4 * npairs * nf * nf * ne if with_k else 0
    rel_ = 4 * ndof * ne if with_r else 0
    out = {
        "element": 4 * nnpe * ne + 2 * 4 * nd * nn + 8 * nf * nn + rel_ + kel + state,
        "reduce": 16 * nn + (4 + vstride) * ne * nnpe + kel + rel_
        + (8 * asm.nnz if with_k else 0) + (8 * nf * nn if with_r else 0),
    }
    # the fused, atomic and coloured designs do the whole step: their algorithmic bytes are the step's compulsory bytes
    name = getattr(wl, "name", None)
    if with_r and with_k and name in STEP_BYTES:
        out["fused"] = out["atomic"] = out["coloured"] = STEP_BYTES[name] * ne
    return out


def pmc_traffic(workload, kernel):
    """HBM bytes per launch (read + written) of `kernel` at `workload`'s full size, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary kept in profiles/ (scaled by a copy-kernel calibration there), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 5 and f[0] == "TRAFFIC" and f[1] == workload and f[2] == kernel:
            return int(f[3]) + int(f[4])
    return None


def cuda_ms(torch, f, reps=10):
    """Average milliseconds of f() between HIP events on the current stream (the assembler's stream)."""
    f()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        f()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


MLP_FLOPS = {"j2_hex8": 2 * (13 * 128 + 128 * 128 * 2 + 128 * 13), "neohookean_tet10": 2 * (9 * 256 + 256 * 256 + 256 * 15)}
MFMA_BF16_PEAK_TFLOPS = 2500.0


def newton_leg(fec, torch, net, n, steps, tol=1e-4, log_every=True):
    """The J2 workload as a quasi-static solve: `steps` load steps of the seeded macroscopic strain path on n^3 hex8,
    each by Newton + Jacobi-CG to a relative residual of `tol`, once with the plain radial return and once with the
    trained surrogate under SurrogateSchedule (decide, freeze, hand over to the plain model when the surrogate
    iterations stall). Reports Newton and CG iterations, wall time, the solutions' difference, the plastic fraction of
    the final state, and the plain residual + Jacobian assembly timed at that plastic state."""
    params, scaling = net.params(), net.scaling()
    # tol: above the floor that FP32 element arithmetic puts under the relative residual (about 2e-5 at 64^3: interior
    # residuals are cancelling sums of O(1) element contributions)
    st = fec.NewtonSolverSettings(max_iters=15, rel_residual_tol=tol, abs_residual_tol=0.0, abs_increment_tol=0.0)
    res = {}
    for surr in (False, True):
        holder = {}

        def setup(asm):
            if surr:
                nt = fec.Surrogate(asm, fec.surrogate.MLP_J2, seed=0)
                nt.set_params(params)
                nt.set_scaling(*scaling)
                nt.set_tolerance(0.05)
                holder["s"] = fec.solvers.SurrogateSchedule(nt)
                holder["phases"] = []

        def schedule(it, h):
            if "s" in holder:
                holder["s"](it, h)

        def log_step(k, hist):
            if "s" in holder:
                holder["phases"].append(list(holder["s"].phase))
            if log_every:
                log(f"[bench] newton {'surrogate' if surr else 'plain'} step {k}: {len(hist)} iterations, "
                    f"rel {hist[-1].rel_R:.1e}")

        torch.cuda.synchronize()
        t0 = time.time()
        asm, p, Uu, hists, _ = fec.solvers.j2_load_steps(fec, n, steps=steps, newton_settings=st, setup=setup,
                                                          surrogate=surr, schedule=schedule, log=log_step)
        torch.cuda.synchronize()
        wall = time.time() - t0
        r = {"newton_iterations": int(sum(len(h) for h in hists)),
             "per_step": [len(h) for h in hists],
             "cg_iterations": int(sum(rec.cg_iters for h in hists for rec in h)),
             "converged_all": bool(all(h[-1].converged for h in hists)),
             "max_final_rel_residual": float(max(h[-1].rel_R for h in hists)),
             "wall_s": wall}
        if surr:
            ph = holder["phases"]
            r["surrogate_iterations"] = int(sum(x.count("surrogate") for x in ph))
            r["plain_iterations_after_handover"] = int(sum(x.count("plain") for x in ph))
        else:
            ne = n ** 3
            state = np.zeros(ne * 8 * 7, dtype=np.float32)
            import ctypes
            asm.ctx.call("fecx_get_state_new", state.ctypes.data_as(ctypes.c_void_p))
            r["plastic_point_fraction_final"] = float((state.reshape(-1, 7)[:, 6] > 0).mean())
            r["assembly_ms_at_final_state"] = cuda_ms(torch, lambda: fec.assemble_vector_and_stiffness_(asm, Uu, p))
        res["surrogate" if surr else "plain"] = r
        res["_U_" + ("s" if surr else "p")] = Uu.cpu().numpy().copy()
    a, b = res.pop("_U_s"), res.pop("_U_p")
    res["solution_rel_diff"] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    res["workload"] = f"j2_hex8 {n}^3, {steps} load steps, Newton to relative residual {tol:g}, Jacobi-CG"
    return res


def surrogate_leg(fec, torch, name, wl, asm_plain, R_plain, K_plain, steps, newton_n=0):
    """The learned path of the J2 and tet10 neo-Hookean configurations: train the network on the GPU on plain-model
    samples (J2: 2M samples of random macroscopic strain paths; NH: 2M quadrature points of this workload), then
    measure the held-out acceptance and error at the residual-check tolerance 0.05, the batched material update of
    2M points (plain vs surrogate with check and fallback vs the network alone, HIP events), and one residual +
    Jacobian assembly of the whole workload with the surrogate physics against the plain one."""
    import copy
    S, L = fec.surrogate, fec._lib
    j2 = name == "j2_hex8"
    plain_kind, surr_kind = (L.J2, L.J2_SURROGATE) if j2 else (L.NEOHOOKEAN, L.NH_SURROGATE)
    wl_s = copy.copy(wl)
    wl_s.physics = fec.J2Surrogate(*wl.physics.p) if j2 else fec.NeoHookeanSurrogate(*wl.physics.p)
    asm, p = build_assembler(fec, wl_s)
    t0 = time.time()
    if j2:
        X, T = S.j2_generate_samples(asm, 100000, steps=20, max_norm=0.01, seed=1)
        Xv, _ = S.j2_generate_samples(asm, 5000, steps=20, max_norm=0.01, seed=2)
    else:
        H = S.qp_gradients(asm, torch.from_numpy(wl.U).to(asm.device)).reshape(-1, 9)
        g = torch.Generator(device=asm.device)
        g.manual_seed(0)
        perm = torch.randperm(H.shape[0], device=asm.device, generator=g)
        X, Xv = H[perm[:2_000_000]].contiguous(), H[perm[2_000_000:2_200_000]].contiguous()
        T = S.material_update(asm, plain_kind, X)
        del H, perm
    net = fec.Surrogate(asm, S.MLP_J2 if j2 else S.MLP_NH, seed=0)
    net.fit_scaling(X, T)
    losses = []
    for lr, ep in ((2e-3, 15), (1e-3, 15), (3e-4, 15), (1e-4, 10)):
        losses.append(net.train(X, T, epochs=ep, batch=16384, lr=lr, seed=int(lr * 1e6)))
    torch.cuda.synchronize()
    t_train = time.time() - t0
    net.set_tolerance(0.05)
    Yv = S.material_update(asm, surr_kind, Xv)
    fb, nv = net.fallbacks()
    Pv = S.material_update(asm, plain_kind, Xv)
    ns = 6 if j2 else 9
    stress_err = ((Yv[:, :ns] - Pv[:, :ns]).abs().max() / Pv[:, :ns].abs().max()).item()
    nb = X.shape[0]
    t_plain = cuda_ms(torch, lambda: S.material_update(asm, plain_kind, X))
    t_surr = cuda_ms(torch, lambda: S.material_update(asm, surr_kind, X))
    fb_b = net.fallbacks()[0]
    t_net = cuda_ms(torch, lambda: net.forward(X))
    tflops = MLP_FLOPS[name] * nb / (t_net * 1e-3) / 1e12
    log(f"[bench] surrogate: trained {t_train:.1f}s, held-out fallback {fb / nv:.3f}, batched update 2M points "
        f"plain {t_plain:.3f} ms, surrogate {t_surr:.3f} ms, network {t_net:.3f} ms ({tflops:.1f} TFLOP/s)")
    # the whole workload with the surrogate physics
    Uu = torch.from_numpy(wl.U[asm.dof.unknown_dofs - 1]).to(asm.device)
    ms = cuda_ms(torch, lambda: fec.assemble_vector_and_stiffness_(asm, Uu, p), reps=max(3, min(steps, 10)))
    fb_a, np_a = net.fallbacks()
    R = fec.residual(asm).cpu().numpy()
    K = asm.stiffness_storage.cpu().numpy()
    out = {
        "network": "13-128-128-128-13" if j2 else "9-256-256-15", "activation": "GELU", "arithmetic": "bf16 MFMA, fp32 accumulation",
        "train_samples": int(nb), "train_s": t_train, "epochs": int(sum(len(x) for x in losses)),
        "final_loss": float(losses[-1][-1]), "check_tolerance": 0.05,
        "heldout": {"samples": int(nv), "fallback_fraction": fb / nv, "max_stress_error_rel": stress_err},
        "batched_update_ms_2M": {"plain": t_plain, "surrogate": t_surr, "network_only": t_net,
                                 "surrogate_fallback_fraction": fb_b / nb},
        "throughput_ratio_vs_plain": t_plain / t_surr,
        "mfma": {"achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / MFMA_BF16_PEAK_TFLOPS,
                 "measured_gemm_peak": gemm_peak_tflops(torch),
                 "in_material_kernel": MLP_FLOPS[name] * nb / (t_surr * 1e-3) / 1e12},
        "assembly_ms_per_step": {"plain_step_time_above": None, "surrogate": ms},
        "assembly_fallback_fraction": fb_a / max(np_a, 1),
        "assembly_rel_diff_vs_plain": {"residual": rel(R, R_plain), "jacobian": rel(K, K_plain)},
    }
    # what the learned path could reach at best: the network's flops at the BF16 MFMA spec and at the measured GEMM
    # peak, against the plain batched update (the network also moves the same bytes, so this is an upper bound)
    peak_ms = MLP_FLOPS[name] * nb / (MFMA_BF16_PEAK_TFLOPS * 1e12) * 1e3
    gemm_ms = MLP_FLOPS[name] * nb / (out["mfma"]["measured_gemm_peak"] * 1e12) * 1e3
    out["bound"] = {"network_flops_per_point": MLP_FLOPS[name], "network_ms_2M_at_spec_peak": peak_ms,
                    "network_ms_2M_at_measured_gemm_peak": gemm_ms, "plain_ms_2M": t_plain,
                    "max_ratio_vs_plain_at_spec_peak": t_plain / peak_ms,
                    "max_ratio_vs_plain_at_measured_gemm_peak": t_plain / gemm_ms}
    if j2 and newton_n:
        out["newton"] = newton_leg(fec, torch, net, newton_n, 20)
    return out


def operator_timings(fec, torch, name, wl, asm, p, Uu):
    """The other entry points on the same workload (HIP events around whole calls, DOF plumbing included): residual
    alone, matrix-free action, assembled action; elements/s and GB/s against the algorithmic bytes per element."""
    Vu = torch.cos(torch.arange(Uu.numel(), dtype=torch.float64, device=Uu.device))
    ne = wl.num_elements
    ops = {"residual": (lambda: fec.assemble_vector_(asm, fec.residual, Uu, p), RESIDUAL_BYTES[name]),
           "matrix_free_action": (lambda: fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p),
                                  MF_ACTION_BYTES[name]),
           "assembled_action": (lambda: fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p),
                                MF_ACTION_BYTES[name])}
    out = {}
    for k, (f, b) in ops.items():
        ms = cuda_ms(torch, f, reps=10)
        fec.reset_profile(asm)
        fec.set_profiling(asm, True)
        for _ in range(5):
            f()
        fec.set_profiling(asm, False)
        kern = {kk: v[1] / v[0] for kk, v in fec.profile(asm).items()}
        out[k] = {"ms": ms, "elements_per_s": ne / (ms * 1e-3), "bytes_per_element": b,
                  "gbs": b * ne / (ms * 1e-3) / 1e9, "kernel_ms": kern}
    log("[bench] operators: " + ", ".join(
        f"{k} {v['ms']:.3f} ms ({v['gbs']:.0f} GB/s; " + " ".join(f"{kk} {t:.3f}" for kk, t in v["kernel_ms"].items())
        + ")" for k, v in out.items()))
    return out


def hbm_ceiling(fec, torch, asm, gib=4, reps=10):
    """The device's HBM ceilings measured by the library's hand-written streaming kernels (fecx_measure_bandwidth) on a
    `gib` GiB buffer: read, write, and copy (bytes read + bytes written), GB/s."""
    import ctypes
    nbytes = gib << 30
    buf = torch.empty(nbytes // 8, dtype=torch.float64, device=asm.device)
    buf.fill_(1.0)
    out = {}
    for op, name in enumerate(("read", "write", "copy")):
        ms = ctypes.c_double(0.0)
        asm.ctx.call("fecx_measure_bandwidth", ctypes.c_int(op), ctypes.c_void_p(buf.data_ptr()),
                     ctypes.c_int64(nbytes), ctypes.c_int(reps), ctypes.byref(ms))
        out[name] = nbytes / (ms.value * 1e-3) / 1e9
    del buf
    torch.cuda.empty_cache()
    return out


def gemm_peak_tflops(torch, n=8192, reps=10):
    """Measured BF16 dense GEMM throughput (torch.matmul on contiguous n x n operands): the attainable MFMA figure
    next to the 2.5 PFLOP/s spec."""
    a = torch.randn(n, n, device="cuda", dtype=torch.bfloat16)
    b = torch.randn(n, n, device="cuda", dtype=torch.bfloat16)
    ms = cuda_ms(torch, lambda: torch.matmul(a, b), reps)
    return 2.0 * n ** 3 / (ms * 1e-3) / 1e12


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def cpu_share():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS when that is set (the share
    of the host a job is given), and the affinity count itself."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    n = min(aff, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else aff
    return n, aff


def oracle_step(orc, wl, pat, nu, split=None):
    """One residual + CSR Jacobian on the CPU with the reference's algorithm: the residual loop, the element stiffness
    storage, then sparse! over the precomputed pattern (built once, as update_dofs! does). `split`, if given, receives
    the three times in seconds."""
    Is, Js, slots = pat
    m = wl.mesh
    kind = {"QUAD4": orc.QUAD4, "HEX8": orc.HEX8, "TET10": orc.TET10}[m.element_type]
    props = wl.physics.props()
    st = np.zeros(m.num_elements * 8 * 7) if wl.physics.kind == orc.J2 else None
    t0 = time.perf_counter()
    R = orc.assemble_residual(kind, m.coords, m.conn, wl.nf, wl.physics.kind, props, wl.U, st,
                              None if st is None else np.zeros_like(st))
    t1 = time.perf_counter()
    storage = orc.assemble_stiffness_storage(kind, m.coords, m.conn, wl.nf, wl.physics.kind, props, wl.U, st)
    t2 = time.perf_counter()
    _, _, vals, _ = orc.sparse_csr(nu, Is, Js, orc.gather_slots(slots, storage))
    t3 = time.perf_counter()
    if split is not None:
        split.append((t1 - t0, t2 - t1, t3 - t2))
    return R, vals


def cpu_baseline_leg(fec, orc, name, n_sample, threads):
    """The oracle's CPU throughput on a bounded sample of the workload (median of 5 steps after 1 warm-up)."""
    wl = fec.workloads.CONFIGS[name](n_sample)
    N = wl.nf * wl.mesh.num_nodes
    unknown, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    orc.set_threads(threads)
    Is, Js, slots, perm = orc.pattern(wl.nf, wl.mesh.conn, d2u)  # setup (update_dofs!), not timed
    pat = (Is, Js, slots)
    oracle_step(orc, wl, pat, len(unknown))  # warm-up
    split = []
    for _ in range(5):
        oracle_step(orc, wl, pat, len(unknown), split)
    split = np.array(split)
    t = float(np.median(split.sum(axis=1)))
    parts = {k: float(np.median(split[:, i])) * 1e3 for i, k in enumerate(("residual", "element_stiffness", "sparse_csr"))}
    return wl.mesh.num_elements, t, parts


def fullsize_checks(fec, orc, torch, wl, asm, p, Uu, threads):
    """Every output of the workload at its full size against the FP64 oracle on the host: the residual, the CSR
    values (oracle element storage reduced with sparse!'s sums, row by row), the matrix-free and the assembled Jacobian
    actions. Relative errors in the 2-norm and the max-norm, the CSR pattern compared bit for bit."""
    m = wl.mesh
    kind = {"QUAD4": orc.QUAD4, "HEX8": orc.HEX8, "TET10": orc.TET10}[m.element_type]
    props = wl.physics.props()
    st = np.zeros(m.num_elements * 8 * 7) if wl.physics.kind == orc.J2 else None
    orc.set_threads(threads)
    out = {}
    t0 = time.time()
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R_gpu = asm.residual_storage.cpu().numpy()
    K_gpu = asm.stiffness_storage.cpu().numpy()
    R = orc.assemble_residual(kind, m.coords, m.conn, wl.nf, wl.physics.kind, props, wl.U, st,
                              None if st is None else np.zeros_like(st))
    out["residual"] = {"l2": rel(R_gpu, R), "max": rel_max(R_gpu, R)}
    del R
    nu = len(asm.dof.unknown_dofs)
    _, d2u = orc.update_dofs(wl.nf * m.num_nodes, asm.dof.dirichlet_dofs)
    storage = orc.assemble_stiffness_storage(kind, m.coords, m.conn, wl.nf, wl.physics.kind, props, wl.U, st)
    rp, cv, vals = orc.sparse_csr_rows(wl.nf, m.conn, d2u, nu, storage)
    del storage
    rp_g, cv_g = (t.cpu().numpy() + 1 for t in asm.pattern())
    out["pattern_bit_exact"] = bool(np.array_equal(rp_g, rp) and np.array_equal(cv_g, cv))
    del rp, cv, rp_g, cv_g
    out["jacobian"] = {"l2": rel(K_gpu, vals), "max": rel_max(K_gpu, vals)}
    del vals, K_gpu
    log(f"[bench] full-size check: residual and Jacobian done ({time.time() - t0:.1f}s)")
    N = wl.nf * m.num_nodes
    Vf = np.cos(np.arange(N, dtype=np.float64))
    Vu = torch.from_numpy(Vf[asm.dof.unknown_dofs - 1]).to(asm.device)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    mf_gpu = asm.stiffness_action_storage.cpu().numpy()
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p)
    asm_gpu = asm.stiffness_action_storage.cpu().numpy()
    for key, g, mf in (("matrix_free_action", mf_gpu, True), ("assembled_action", asm_gpu, False)):
        ref = orc.assemble_action(kind, m.coords, m.conn, wl.nf, wl.physics.kind, props, wl.U, Vf, st, matrix_free=mf)
        out[key] = {"l2": rel(g, ref), "max": rel_max(g, ref)}
    out["against"] = (f"FP64 oracle (C++ restatement of the reference algorithm) at the full size, "
                      f"{m.num_elements} elements; V = cos(dof index) for the actions")
    log(f"[bench] full-size check done ({time.time() - t0:.1f}s): " + ", ".join(
        f"{k} {v['l2']:.2e}/{v['max']:.2e}" for k, v in out.items() if isinstance(v, dict)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 20 warm-up steps bring the device from the idle of the set-up to its steady clocks (neo-Hookean hex8 96^3: 0.972
    # ms/step timed after 3 warm-up steps, 0.927 after 20 or 50); 100 timed steps are ~0.1 s at the headline size
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="neohookean_hex8", choices=sorted(SIZES))
    ap.add_argument("--n", type=int, default=None, help="elements per direction (default: the config's size)")
    ap.add_argument("--chunks", type=int, default=0, help="pipelined pieces per matrix assembly (0 or 1: one piece)")
    ap.add_argument("--scatter", default="auto", choices=["fused", "two_phase", "atomic", "coloured", "auto"],
                    help="scatter design of the matrix assembly (auto: the library's default, fused on structured "
                         "hex8 meshes with vector-valued materials, two-phase elsewhere)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the full-size parity and CPU legs (profiling runs)")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate leg (J2, tet10 NH)")
    ap.add_argument("--no-operators", action="store_true", help="skip the timings of the other entry points")
    ap.add_argument("--fp64-geometry", action="store_true",
                    help="tet10: element geometry, field gradient and stress in FP64 (fecx_set_fp64_geometry)")
    ap.add_argument("--newton", type=int, default=0, metavar="N",
                    help="j2_hex8 learned leg: also solve the 20 load steps on N^3 elements by Newton, plain and "
                         "with the surrogate (N = 64: the configuration's full size)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (HBM counter calibration)")
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="only measure the HBM read / write / copy ceilings with the library's streaming kernels")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 8, "calibrate_copy_bytes_written": n * 8}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): assembly across GPUs (element "
              "partition, reverse halo exchange of partial rows) is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch
    fec = ge.load_package()
    if args.hbm_ceiling:
        wl = fec.workloads.poisson_hex8(2)
        asm, _ = build_assembler(fec, wl)
        print(json.dumps({"hbm_ceiling_gbs": hbm_ceiling(fec, torch, asm), "buffer_gib": 4, "reps": 10}), flush=True)
        return
    name = args.workload
    n_full, n_sample = SIZES[name]
    n = args.n or n_full
    log(f"[bench] building {name} n={n}")
    t0 = time.time()
    wl = fec.workloads.CONFIGS[name](n)
    asm, p = build_assembler(fec, wl)
    Uu = torch.from_numpy(wl.U[asm.dof.unknown_dofs - 1]).to(asm.device)
    fec.set_chunks(asm, args.chunks)
    import ctypes
    design = ["fused", "two_phase", "atomic", "coloured", "auto"].index(args.scatter)
    asm.ctx.call("fecx_set_scatter", ctypes.c_int(design))
    asm.ctx.call("fecx_set_fp64_geometry", ctypes.c_int(1 if args.fp64_geometry else 0))
    fp64_geo = args.fp64_geometry and wl.mesh.element_type == "TET10"
    path = ctypes.c_int(-1)
    asm.ctx.call("fecx_get_scatter_path", ctypes.byref(path))
    scatter_path = ["fused", "two_phase", "atomic", "coloured"][path.value]
    log(f"[bench] setup {time.time() - t0:.1f}s  NE={wl.num_elements} nnz={asm.nnz}  scatter {scatter_path}")

    for _ in range(args.warmup):
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"[bench] {dt * 1e3:.3f} ms/step")

    # per-kernel durations with HIP events on the stream each kernel runs on
    fec.reset_profile(asm)
    fec.set_profiling(asm, True)
    n_prof = max(3, min(args.steps, 10))
    for _ in range(n_prof):
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
    fec.set_profiling(asm, False)
    prof = fec.profile(asm)
    kb = kernel_bytes(wl, asm)
    # per step: a matrix assembly may launch each kernel once per pipelined piece
    kernels = {k: {"avg_ms": v[1] / v[0], "launches_per_step": v[0] / n_prof, "ms_per_step": v[1] / n_prof,
                   "bytes": kb.get(k)} for k, v in prof.items()}
    for k, v in kernels.items():
        if v["bytes"]:
            v["gbs"] = v["bytes"] / (v["ms_per_step"] * 1e-3) / 1e9
        tr = pmc_traffic(name, k) if n == n_full else None
        if tr:
            v["traffic"] = tr
    log("[bench] kernels: " + ", ".join(f"{k} {v['ms_per_step']:.3f} ms ({v['launches_per_step']:.0f} launches)"
                                        for k, v in kernels.items()))
    # The roofline of the step's assembly kernels together (the DOF update excluded), priced on the compulsory bytes
    # per element of residual + tangent: int32 connectivity, FP32 coordinates, FP64 u, FP64 outputs written once, the
    # int32 COO->CSR map (NDOF^2 per element) and FP32 state.
    STEP_KERNELS = ("element", "reduce", "fused", "atomic", "coloured")
    kern_ms = sum(v["ms_per_step"] for k, v in kernels.items() if k in STEP_KERNELS)
    bpe = STEP_BYTES.get(name)
    traffic = None
    if n == n_full and all("traffic" in v for k, v in kernels.items() if k in STEP_KERNELS):
        traffic = sum(v["traffic"] for k, v in kernels.items() if k in STEP_KERNELS)
    out = {
        "metric": METRIC,
        "value": wl.num_elements / dt,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{name} {n}^3" if wl.mesh.num_dims == 3 else f"{name} {n}^2",
                   "elements": int(wl.num_elements), "nnz": int(asm.nnz), "step": "residual + CSR Jacobian",
                   "element_math": "fp32, geometry and stress fp64" if fp64_geo else "fp32", "reductions": "fp64",
                   "chunks": max(1, args.chunks),
                   "scatter": scatter_path},
        "kernels": kernels,
    }
    if bpe:
        ach = bpe * wl.num_elements / (kern_ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "+".join(k for k in kernels if k in STEP_KERNELS),
                           "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                           "traffic": traffic, "bytes_per_element": bpe, "kernel_ms_per_step": kern_ms,
                           "measured_ceiling_gbs": hbm_ceiling(fec, torch, asm, gib=2, reps=5)}
        eff = bpe * wl.num_elements / dt / 1e9
        out["step_roofline"] = {"bytes_per_element": bpe, "achieved": eff, "frac": eff / HBM_PEAK_GBS,
                                "basis": "wall time per step (DOF update and host synchronisation included)"}
        if not args.no_operators:
            out["operators"] = operator_timings(fec, torch, name, wl, asm, p, Uu)

    if name in MLP_FLOPS and not args.no_surrogate:
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
        R_plain = fec.residual(asm).cpu().numpy().copy()
        K_plain = asm.stiffness_storage.cpu().numpy().copy()
        out["surrogate"] = surrogate_leg(fec, torch, name, wl, asm, R_plain, K_plain, args.steps, args.newton)
        out["surrogate"]["assembly_ms_per_step"]["plain_step_time_above"] = dt * 1e3

    if not args.no_check:
        orc = ge.load_oracle()
        threads, aff = cpu_share()
        out["rel_err"] = fullsize_checks(fec, orc, torch, wl, asm, p, Uu, threads)
        if not args.no_cpu_baseline:
            ne_s, t_cpu, parts = cpu_baseline_leg(fec, orc, name, n_sample, threads)
            out["cpu_baseline"] = {"value": ne_s / t_cpu, "unit": "elements/s", "cores": threads, "kind": "port",
                                   "sample": f"{name} n={n_sample}: {ne_s} elements, residual + element stiffness + "
                                             f"sparse! COO->CSR over the pattern built at setup, median of 5 after "
                                             f"1 warm-up (C++ FP64 restatement of the reference algorithm, OpenMP on "
                                             f"{threads} threads: the host share given to this job, affinity {aff})",
                                   "split_ms": parts}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

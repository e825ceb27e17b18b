"""Benchmark: symmetric basis construction + H x on the 6x6 J1-J2 model.

Workload (one MI355X): J1-J2 Heisenberg model on the 6x6 periodic square
lattice, J1 = 1, J2 = 0.5 (Pauli matrices per bond as in
chapel/data/heisenberg_square_6x6.yaml), Hamming weight 18, translations x
C4v x spin inversion (576 elements), 15 804 956 representatives.

A step is one product y = H x over the whole basis with x resident in HBM
(seeded N(0,1) FP64), from the CSR cache built on the GPU.  The line also
reports the GPU basis build (states/s), the FP32 block-of-8 product, the
on-the-fly fused product, and the CPU baseline (the oracle, a C/OpenMP
restatement of the reference's CPU algorithm, on a bounded sample).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload chain-sectors   (56 sectors of the 28-site chain)
       python bench.py --workload configs         (16- and 32-site chains, 4x4 Hubbard)
       python bench.py --workload gather-ceiling  (the FP64 kernel on synthetic gather windows)
       python bench.py --workload order-model     (L2 misses of other orders of the 6x6 matrix)
       python bench.py --workload column-distance (how far the 6x6 product's columns lie from their rows)
       python bench.py --workload copy-calibration (a copy of known size for the counters)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, os.path.join(REPO, "lattice-symmetries_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_KIND = ("L2 (TCC) miss bytes per launch: rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy of known "
                "size; what leaves the L2 for the fabric (Infinity Cache or HBM), an upper bound on HBM bytes")


def traffic_rate(traffic, kernel_ms):
    """The measured L2-miss bytes per launch over the measured kernel time: the
    rate the memory system actually sustains, beside the algorithmic one."""
    if traffic is None or not kernel_ms:
        return {"traffic_kind": TRAFFIC_KIND, "traffic_GBs": None, "traffic_frac": None}
    gbs = traffic / (kernel_ms * 1e-3) / 1e9
    return {"traffic_kind": TRAFFIC_KIND, "traffic_GBs": gbs, "traffic_frac": gbs / HBM_PEAK_GBS}


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` measured by rocprofv3 (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy of known size), as recorded in
    profiles/traffic_per_launch.json; None when this kernel was not profiled."""
    try:
        with open(os.path.join(REPO, "profiles", "traffic_per_launch.json")) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None, None
    e = table.get(workload, {}).get(kernel)
    return (e["bytes"], e["source"]) if e else (None, None)


def workload(L=6, J2=0.5):
    import lsamd
    from lsamd import models
    n = L * L
    gens = models.square_generators(L)
    S = lsamd.Symmetries([lsamd.Symmetry(g, 0) for g in gens])
    expr = models.j1j2_square(L, 1.0, J2)
    return n, n // 2, S, expr


def cpu_baseline(S, n, h, reps, expr, seconds):
    """The oracle (a C/OpenMP restatement of the reference's CPU algorithm) on
    every host core this process may use, on bounded samples of the workload:
    - the push-form product (Benes orbit search per beta, binary-search
      index, atomic accumulate) from source rows spread evenly over the whole
      basis (every stride-th representative);
    - the basis enumeration (Gosper steps + orbit filter) over eight candidate
      ranges spread evenly over the combinadics rank range."""
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    g = O.Group(n, S.elements(), S.characters.real, S.characters.imag, 1, masks=S.masks, shifts=S.shifts)
    g.use_benes(True)
    diag, off = expr.split()
    d, o = O.Terms(diag), O.Terms(off)
    x = np.random.default_rng(42).standard_normal(reps.size)
    y = np.zeros(reps.size)
    stride = max(1, reps.size // 4000)
    t0 = time.perf_counter()
    O.matvec_push_parallel(g, reps, d, o, x, y, 0, reps.size, threads, stride=stride)
    dt = time.perf_counter() - t0
    per_row = dt / ((reps.size + stride - 1) // stride)
    stride = max(1, int(per_row * reps.size / (0.7 * seconds)))
    y[:] = 0
    t0 = time.perf_counter()
    nnz = O.matvec_push_parallel(g, reps, d, o, x, y, 0, reps.size, threads, stride=stride)
    dt = time.perf_counter() - t0
    rows = (reps.size + stride - 1) // stride
    # enumeration: eight equal candidate ranges spread over the rank range
    lo = (1 << h) - 1
    hi = lo << (n - h - 1)
    r0, r1 = O.state_to_index(lo), O.state_to_index(hi)
    total = r1 - r0 + 1
    chunk = min(1 << 20, total)
    t0 = time.perf_counter()
    g.count_parallel(O.index_to_state(r0, h), chunk, threads)
    t_probe = time.perf_counter() - t0
    parts = 8 if total >= 8 else 1
    chunk = int(min(total // parts, max(1 << 16, chunk * 0.3 * seconds / parts / max(t_probe, 1e-3))))
    found = 0
    t0 = time.perf_counter()
    for q in range(parts):
        start = r0 + (total - chunk) * q // max(parts - 1, 1)
        found += g.count_parallel(O.index_to_state(start, h), chunk, threads)
    t_enum = time.perf_counter() - t0
    return {"value": nnz / dt, "unit": "nonzeros/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"push-form H x from every {stride}-th of {reps.size} representatives ({rows} source rows, "
                      f"{nnz} nonzeros, {dt:.1f} s, Benes orbit search, {threads} OpenMP threads)",
            "basis": {"candidates_per_s": parts * chunk / t_enum, "states_per_s": found / t_enum,
                      "sample": f"enumeration of {parts} ranges of {chunk} candidates spread over the {total} "
                                f"candidates ({found} representatives, {t_enum:.1f} s, {threads} OpenMP threads)"}}


def sectors_bench(args):
    """All 56 sectors of the 28-site J1-J2 chain (k = 0..27 x spin inversion
    +-1, J2 = 0.5, Hamming weight 14) as one block-diagonal complex operator:
    the FP32 block-of-8 product as one persistent graph-replayed launch, then
    the batched block Lanczos (lowest 4 eigenvalues per sector), checked
    against FP64 single-sector Lanczos on a few sectors and against the exact
    Majumdar-Ghosh ground-state energy -3N/2 at k = 0 and k = pi."""
    import lsamd
    from lsamd import models, sectors as LS
    lib = lsamd.load()
    n, h = args.chain, args.chain // 2
    expr = models.j1j2_chain(n, 1.0, 0.5)
    t0 = time.perf_counter()
    batch, bases, ops = LS.chain_sector_batch(n, h, expr, keep_operators=True)
    lib.ls_amd_synchronize()
    t_setup = time.perf_counter() - t0
    dim, nnz_off = batch.number_states, batch.nnz
    log(f"sectors: {batch.count} sectors, {dim} states, {nnz_off} off-diagonal nonzeros, set-up {t_setup:.1f} s")
    names = [f"{k}{'+' if si == 1 else '-'}" for k, si in batch.sectors]
    keep = set(names) if args.reference_sectors == "all" else set(args.reference_sectors.split(","))
    ref_ops = {nm: op for nm, op in zip(names, ops) if nm in keep}
    del bases, ops
    sched = batch.autotune(8)
    log(f"autotune: {sched}")
    rng = np.random.default_rng(42)
    X = rng.standard_normal((dim, 2, 8)).astype(np.float32)
    dx = lsamd.DeviceArray.from_numpy(X)
    dy = lsamd.DeviceArray(X.nbytes)
    for _ in range(args.warmup):
        batch.matvec_device(8, dx, dy, 32)
    lib.ls_amd_synchronize()
    k_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.matvec_device(8, dx, dy, 32)
        k_ms.append(lib.ls_amd_last_kernel_ms(6))
    lib.ls_amd_synchronize()
    t_steps = time.perf_counter() - t0
    ms = float(np.mean(k_ms))
    nnz = nnz_off + dim
    # compulsory bytes: the CSR stream (column 4 + complex FP32 value 8 per
    # nonzero), row pointer 8 and diagonal 4 per row, x read once and y
    # written once (8 complex FP32 columns, 64 B each per row)
    bytes_per_launch = 12.0 * nnz_off + 140.0 * dim
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9
    # the per-nonzero model that charges every x-row gather (76 B per nonzero)
    gather_model = (76.0 * nnz_off + 140.0 * dim) / (ms * 1e-3) / 1e9
    sector_kernel = (f"sector_spmv_kernel<8, {sched['lanes']}, float, float>"
                     f"{' die-aware' if sched['die_aware'] else ''} x{sched['blocks']}")
    traffic, traffic_src = profiled_traffic(f"chain{n}_sectors", sector_kernel)
    value = 8.0 * nnz * args.steps / t_steps
    del dx, dy
    sched4 = batch.autotune(4)  # the eigensolver's products carry 4 complex columns
    log(f"autotune (4 columns): {sched4}")
    r = batch.eigensolve(nev=4, block_size=4, max_blocks=args.sector_blocks, keep_blocks=args.sector_keep,
                         max_cycles=args.sector_cycles, tol=1e-7, switch_tol=args.sector_switch_tol, seed=1)
    log(f"eigensolve: {r['cycles']} cycles ({r['fp64_cycles']} FP64), {r['products']} products, "
        f"{r['total_seconds']:.1f} s, max residual {r['max_residual']:.2e}")
    ev = r["eigenvalues"]
    checks = {}
    for nm, op in ref_ops.items():
        s = names.index(nm)
        ref = op.lanczos("otf_complex", max_steps=args.reference_steps, k=4, tol=1e-12)
        e_ref = np.array(ref["eigenvalues"][:4])
        d = float(np.abs(ev[s] - e_ref).max() / abs(e_ref[0]))
        checks[nm] = {"batched": [float(v) for v in ev[s]], "fp64_single_sector": [float(v) for v in e_ref],
                      "max_rel_diff": d, "reference_steps": ref["steps"]}
        log(f"sector {nm}: batched {ev[s]} vs FP64 {e_ref}: {d:.2e}")
    mg = -1.5 * n
    s0, spi = names.index("0+"), names.index(f"{n // 2}+")
    line = {
        "metric": "sector-batched H·x (complex, block of 8) nonzeros x RHS/s + block-Lanczos per sector",
        "value": value,
        "unit": "nonzero-RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_steps / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded N(0,1) FP32 right-hand sides (seed 42)",
        "config": {"workload": f"J1-J2 chain N={n}, J2=0.5, Hamming weight {h}, {batch.count} sectors "
                               "(k x spin inversion), block of 8 complex RHS, one graph-replayed launch",
                   "states": dim, "off_diagonal_nonzeros": nnz_off, "schedule": sched},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     **traffic_rate(traffic, ms),
                     "kernel": sector_kernel, "kernel_ms": ms, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "bytes_model": "compulsory: CSR stream once, x read once, y written once",
                     "gather_model_GBs": gather_model, "gather_model_frac": gather_model / HBM_PEAK_GBS},
        "setup_seconds": t_setup,
        "eigensolve": {"method": "complex thick-restart block Lanczos, FP32 then FP64",
                       "block_complex_columns": 4, "schedule": sched4, "max_blocks": args.sector_blocks,
                       "keep_blocks": args.sector_keep, "switch_tol": args.sector_switch_tol,
                       "cycles": r["cycles"], "fp64_cycles": r["fp64_cycles"], "products": r["products"],
                       "product_seconds": r["product_seconds"], "seconds": r["total_seconds"],
                       "max_residual": r["max_residual"], "converged": r["converged"],
                       "E0_k0": float(ev[s0][0]), "E0_kpi": float(ev[spi][0]), "E0_exact": mg,
                       "E0_rel_err": float(max(abs(ev[s0][0] - mg), abs(ev[spi][0] - mg)) / abs(mg)),
                       "lowest_per_sector": {nm: [float(v) for v in ev[i]] for i, nm in enumerate(names)}},
        "fp64_reference": checks,
    }
    print(json.dumps(line), flush=True)


def sched_suffix(sched, packed):
    """Template flags of an FP64 product kernel: non-temporal loads, die-aware
    mapping, value table (the packed format)."""
    f = sched["flags_f64"]
    tf = lambda b: "true" if b else "false"  # noqa: E731
    return f"{tf(f & 1)}, {tf(f & 2)}, {tf(packed)}"


def timed(lib, family, steps, launch):
    """Mean engine-timer duration (ms) of `launch` over `steps` launches after one warm-up."""
    launch()
    lib.ls_amd_synchronize()
    ms = []
    for _ in range(steps):
        launch()
        ms.append(lib.ls_amd_last_kernel_ms(family))
    lib.ls_amd_synchronize()
    return float(np.mean(ms))


def fp64_product(lib, op, dim, steps, x, y):
    """Autotunes and times the FP64 CSR product of `op` (built) against the
    per-nonzero model of the headline (20 B per stored nonzero, 32 B per row)."""
    op.autotune()
    sched = op.schedule
    info = op.csr_info
    ms = timed(lib, 0, steps, lambda: op.csr_matvec(x, y))
    nnz_off = info["nnz"]
    achieved = (20.0 * nnz_off + 32.0 * dim) / (ms * 1e-3) / 1e9
    if info["wide"]:
        f = sched["flags_f64"]
        kernel = (f"csr_f64_wide_kernel<{sched['lanes_f64']}, {'true' if f & 1 else 'false'}, "
                  f"{'true' if f & 2 else 'false'}>")
    else:
        kernel = f"csr_f64_kernel<{sched['lanes_f64']}, {sched_suffix(sched, info['packed'])}>"
    if sched["flags_f64"] & 6 == 6:
        kernel += " full grid"
    table = "packed" if info["packed"] else ("wide" if info["wide"] else None)
    return {"kernel": kernel, "kernel_ms": ms, "off_diagonal_nonzeros": nnz_off, "packed": info["packed"],
            "table": table, "value_table": info["table_size"] if table else None, "schedule": sched,
            "nonzeros_per_s": (nnz_off + dim) / (ms * 1e-3), "achieved_GBs": achieved,
            "frac": achieved / HBM_PEAK_GBS}


def hdf5_io(basis, x, dim):
    """HDF5 interchange of the basis and one vector between device memory and
    a file in the reference's layout (basis/representatives u64 [dim],
    hamiltonian/eigenvectors f64 [1][dim]), through liblsamd_hdf5.so.  The
    file is written to the temporary directory and read back at once, so the
    reads come from the page cache; nothing is fsync'ed."""
    import tempfile

    import lsamd
    from lsamd import hdf5 as H
    if not H.available():
        return {"available": False}
    path = os.path.join(tempfile.gettempdir(), f"lsamd_bench_{os.getpid()}.h5")
    out = {"available": True, "file": "basis/representatives u64 [dim] + hamiltonian/eigenvectors f64 [1][dim]"}
    try:
        H.make_file(path)
        for g in ("basis", "hamiltonian"):
            H.make_group(path, g)
        t0 = time.perf_counter()
        H.write_dataset_as_blocks(path, "basis/representatives", basis.states)
        t_w = time.perf_counter() - t0
        t0 = time.perf_counter()
        states = H.read_dataset_as_blocks(path, "basis/representatives", rank=1, dtype=np.uint64)
        lsamd.DeviceArray.from_numpy(states)
        t_r = time.perf_counter() - t0
        t0 = time.perf_counter()
        H.write_device_dataset(path, "hamiltonian/eigenvectors", x, np.float64, (1, dim))
        t_wv = time.perf_counter() - t0
        t0 = time.perf_counter()
        dv, _ = H.read_device_dataset(path, "hamiltonian/eigenvectors", np.float64, 2)
        t_rv = time.perf_counter() - t0
        ok = bool(np.array_equal(dv.to_numpy(np.float64, (dim,)), x.to_numpy(np.float64, (dim,)))
                  and np.array_equal(states, basis.states))
        gb = 8 * dim / 1e9
        out.update({"representatives_write_GBs": gb / t_w, "representatives_read_GBs": gb / t_r,
                    "vector_write_GBs": gb / t_wv, "vector_read_GBs": gb / t_rv, "bytes_each": 8 * dim,
                    "round_trip_exact": ok,
                    "note": "device -> file and file -> device, PCIe copies included; reads from the page cache"})
        log(f"HDF5: representatives {gb / t_w:.2f} GB/s write, {gb / t_r:.2f} GB/s read; vector "
            f"{gb / t_wv:.2f} GB/s write, {gb / t_rv:.2f} GB/s read (round trip exact: {ok})")
    finally:
        if os.path.exists(path):
            os.remove(path)
    return out


def general_case(lib, basis, L, dim, steps, rng):
    """The FP64 product beyond the 256-value table: the same 6x6 matrix
    stored unpacked (column and FP64 value per entry), and a model with its
    own coupling on every bond (2836 distinct values), stored with the wide
    value table (column and 16-bit value index) and unpacked."""
    import lsamd
    from lsamd import models
    x = lsamd.DeviceArray.from_numpy(rng.standard_normal(dim))
    y = lsamd.DeviceArray(8 * dim)
    out = {}
    op = lsamd.DeviceOperator(basis, models.j1j2_square(L, 1.0, 0.5))
    op.set_csr_packing(False)
    t0 = time.perf_counter()
    op.build_csr(64)
    out["fp64_unpacked"] = {"csr_build_seconds": time.perf_counter() - t0,
                            **fp64_product(lib, op, dim, steps, x, y)}
    del op
    op = lsamd.DeviceOperator(basis, models.random_bond_square(L, seed=7))
    t0 = time.perf_counter()
    op.build_csr(64)
    model = (f"J1-J2 bonds of the {L}x{L} lattice with seeded couplings J (1 + 0.5 u), u in [-1/2, 1/2) per "
             "bond (seed 7), on the same symmetric basis (push-form product of the reference's formula)")
    out["fp64_random_couplings"] = {"model": model, "csr_build_seconds": time.perf_counter() - t0,
                                    **fp64_product(lib, op, dim, steps, x, y)}
    want = lsamd.DeviceArray(8 * dim)
    op.csr_matvec(x, want)
    sw = op.schedule
    del op
    op = lsamd.DeviceOperator(basis, models.random_bond_square(L, seed=7))
    op.set_csr_packing(False)
    t0 = time.perf_counter()
    op.build_csr(64)
    out["fp64_random_couplings_unpacked"] = {"model": model, "csr_build_seconds": time.perf_counter() - t0,
                                             **fp64_product(lib, op, dim, steps, x, y)}
    # the two formats store the same values: with the same lanes per row (the
    # order of each row's sum) the products agree bit for bit
    op.set_schedule(sw["lanes_f64"], sw["flags_f64"], sw["lanes_f32"], sw["flags_f32"])
    op.csr_matvec(x, y)
    out["fp64_random_couplings"]["same_bits_as_unpacked"] = bool(
        np.array_equal(want.to_numpy(np.float64, (dim,)), y.to_numpy(np.float64, (dim,))))
    del op, want
    for k, v in out.items():
        traffic, src = profiled_traffic(f"square{L}_general", v["kernel"])
        v.update({"traffic": traffic, "traffic_source": src, **traffic_rate(traffic, v["kernel_ms"])})
        log(f"{k}: {v['kernel_ms']:.3f} ms, {v['frac']:.3f} of the roofline, value table {v['table']}")
    return out


def configs_bench(args):
    """The three smaller workloads, one JSON line each: the 16-site chain
    (k = 0, parity +, inversion +, S.S couplings, 100-step FP64 Lanczos), the
    32-site chain (k = 0, parity +, inversion +) and the 4x4 Hubbard model
    (t = 1, U = 4, 7 + 7 fermions, T x C4v A1 with fermionic signs, complex
    x).  Each reports its basis build, its CSR build and its products against
    the HBM roofline."""
    import lsamd
    from lsamd import models
    lib = lsamd.load()
    rng = np.random.default_rng(42)

    def spin_chain(n, expr):
        gens = [models.chain_translation(n), models.chain_parity(n)]
        S = lsamd.Symmetries([lsamd.Symmetry(g, 0) for g in gens])
        return lsamd.DeviceBasis(n, n // 2, 1, S), expr

    def hubbard(L):
        n = L * L
        gens = [np.concatenate([g, g + n]).astype(np.int32) for g in models.square_generators(L)]
        S = lsamd.Symmetries([lsamd.Symmetry(g, 0) for g in gens])
        return lsamd.DeviceBasis.fermions(n, 7, 7, S), models.hubbard(n, models.square_nn_bonds(L), t=1.0, U=4.0)

    cases = [("chain16", "Heisenberg chain N=16, H = sum S_i.S_i+1, Hamming weight 8, k=0, parity +, inversion + "
                         "(64 elements)", lambda: spin_chain(16, models.heisenberg_chain(16, 1.0, 0.25))),
             ("chain32", "Heisenberg chain N=32, J=1 (sigma.sigma), Hamming weight 16, k=0, parity +, inversion + "
                         "(128 elements)", lambda: spin_chain(32, models.heisenberg_chain(32))),
             ("hubbard4x4", "Hubbard 4x4, t=1, U=4, 7 up + 7 down, T x C4v A1 with fermionic signs "
                            "(128 elements), complex x", lambda: hubbard(4))]
    for name, desc, make in cases:
        basis, expr = make()
        lib.ls_amd_synchronize()
        t0 = time.perf_counter()
        dim = basis.build()
        lib.ls_amd_synchronize()
        t_build = time.perf_counter() - t0
        op = lsamd.DeviceOperator(basis, expr)
        t0 = time.perf_counter()
        nnz_off = op.build_csr(96)
        t_csr = time.perf_counter() - t0
        op.autotune()
        sched = op.schedule
        packed = op.csr_info["packed"]
        nnz = nnz_off + dim
        line = {"metric": f"{name}: H·x nonzeros/s and basis states/s", "unit": "nonzeros/s", "n_gpus": 1,
                "steps": args.steps, "higher_is_better": True, "dtype": "f64",
                "data": "synthetic: seeded N(0,1) x (seed 42)",
                "config": {"workload": desc, "states": dim, "off_diagonal_nonzeros": nnz_off, "schedule": sched},
                "basis_build": {"seconds": t_build, "states_per_s": dim / t_build,
                                "candidates": basis.number_candidates,
                                "candidates_per_s": basis.number_candidates / t_build},
                "csr_build_seconds": t_csr}
        if name == "hubbard4x4":
            # complex x on the real A1 matrix: the interleaved complex vector
            # is an [n][2] real block, so the FP64 block product with k = 2
            # is the complex product (28 B per nonzero: column, value, x)
            xc = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
            dx = lsamd.DeviceArray.from_numpy(xc)
            dy = lsamd.DeviceArray(xc.nbytes)
            dz = lsamd.DeviceArray(xc.nbytes)
            k_ms = timed(lib, 5, args.steps, lambda: op.csr_matvec_f64_block(2, dx, dy))
            otf_ms = timed(lib, 2, max(2, args.steps // 4), lambda: op.matvec_complex(dx, dz))
            y, z = dy.to_numpy(np.complex128, (dim,)), dz.to_numpy(np.complex128, (dim,))
            line["csr_vs_otf_max_rel_diff"] = float(np.abs(y - z).max() / np.abs(z).max())
            bytes_per_launch = 28.0 * nnz_off + 48.0 * dim
            kernel = f"csr_block_kernel<2, {sched['lanes_f64']}, double, double, {sched_suffix(sched, packed)}> (complex x)"
            line["otf_complex"] = {"kernel_ms": otf_ms, "nonzeros_per_s": nnz / (otf_ms * 1e-3)}
        else:
            x = lsamd.DeviceArray.from_numpy(rng.standard_normal(dim))
            y = lsamd.DeviceArray(8 * dim)
            k_ms = timed(lib, 0, args.steps, lambda: op.csr_matvec(x, y))
            otf_ms = timed(lib, 2, max(2, args.steps // 4), lambda: op.matvec_otf(x, y))
            bytes_per_launch = 20.0 * nnz_off + 32.0 * dim
            kernel = f"csr_f64_kernel<{sched['lanes_f64']}, {sched_suffix(sched, packed)}>"
            line["otf"] = {"kernel_ms": otf_ms, "nonzeros_per_s": nnz / (otf_ms * 1e-3)}
            if name == "chain16":
                r = op.lanczos("fp64", max_steps=100, k=1, tol=0.0)
                line["lanczos_100"] = {"E0": r["eigenvalues"][0], "E0_per_site": r["eigenvalues"][0] / 16,
                                       "steps": r["steps"], "seconds": r["total_seconds"]}
        achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9
        if sched["flags_f64"] & 6 == 6:
            kernel += " full grid"
        traffic, traffic_src = profiled_traffic(name, kernel)
        line.update({"value": nnz / (k_ms * 1e-3), "ms_per_step": k_ms,
                     "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                                  **traffic_rate(traffic, k_ms), "kernel": kernel,
                                  "kernel_ms": k_ms, "algorithmic_bytes_per_launch": bytes_per_launch}})
        if bytes_per_launch < 64e6:
            # a product this small finishes within a few microseconds of its launch
            line["roofline"]["note"] = (f"{bytes_per_launch / 1e3:.0f} KB per launch: the product is bound by launch "
                                        "latency, not memory; the fraction does not rate the kernel")
        log(f"{name}: {dim} states in {t_build:.3f} s, {nnz_off} off-diagonal nonzeros, product {k_ms:.3f} ms "
            f"({achieved / HBM_PEAK_GBS:.2f} of the roofline), on the fly {otf_ms:.3f} ms")
        print(json.dumps(line), flush=True)
        del op, basis


def probes():
    """libprobes.so: measurement kernels built beside the engine (not part of it)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(REPO, "lattice-symmetries_amd", "libprobes.so"))
    lib.lsamd_probe_error.restype = C.c_char_p
    lib.lsamd_probe_gather.argtypes = [C.c_int64, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_double)]
    lib.lsamd_probe_copy.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_double)]
    return lib


def copy_calibration(args):
    """A plain device copy of known size, run under the profiler's FETCH_SIZE /
    WRITE_SIZE counters to scale them to bytes."""
    import ctypes as C
    lib = probes()
    nbytes = args.copy_mib << 20
    ms = C.c_double()
    if lib.lsamd_probe_copy(nbytes, args.steps, C.byref(ms)) != 0:
        raise SystemExit(lib.lsamd_probe_error().decode())
    print(json.dumps({"workload": "copy-calibration", "kernel": "copy_kernel", "bytes_read_per_copy": nbytes,
                      "bytes_written_per_copy": nbytes, "copies": args.steps, "ms_per_copy": ms.value,
                      "GBs": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)


def gather_ceiling(args):
    """The FP64 product kernel on synthetic matrices of the 6x6 workload's
    shape (15 804 956 rows x 74 entries) whose rows gather from windows of W
    columns: small windows keep the gathered vector in L2, W = rows spreads
    the gathers over the whole vector.  The best time at small W bounds what
    any reordering of the real matrix could reach with one gather per
    nonzero."""
    import ctypes as C
    lib = probes()
    rows, per_row = 15804956, 74
    out = []
    for packed in (1, 0):
        for window in (1 << 16, 1 << 18, 1 << 20, 1 << 22, rows):
            for flags in (0, 6):  # plain grid; die-aware mapping, one workgroup per tile
                for lanes in (8, 16, 32):
                    ms = C.c_double()
                    rc = lib.lsamd_probe_gather(rows, per_row, window, lanes, flags, packed, 5, C.byref(ms))
                    if rc != 0:
                        raise SystemExit(lib.lsamd_probe_error().decode())
                    nnz = rows * per_row
                    out.append({"packed": packed, "window_columns": window, "window_MB": window * 8 / 1e6,
                                "flags": flags, "lanes": lanes, "ms": ms.value,
                                "gathers_per_s": nnz / (ms.value * 1e-3),
                                "algorithmic_GBs": (20.0 * nnz + 32.0 * rows) / (ms.value * 1e-3) / 1e9})
                    log(f"packed {packed} window {window:9d} flags {flags} lanes {lanes:2d}: {ms.value:.3f} ms, "
                        f"{nnz / (ms.value * 1e-3) / 1e9:.1f} G gathers/s")
    print(json.dumps({"workload": "gather-ceiling", "rows": rows, "per_row": per_row, "results": out}), flush=True)


def column_distance(args):
    """How far the 6x6 product's columns lie from their row: the distribution
    of |column - row| over every stored entry, and the share of entries that
    a window of W vector entries around a tile of T rows would hold (an
    LDS-resident slice of x; 160 KB of LDS holds 20 480 FP64 values)."""
    import lsamd
    n, h, S, expr = workload(args.lattice)
    basis = lsamd.DeviceBasis(n, h, 1, S)
    dim = basis.build()
    op = lsamd.DeviceOperator(basis, expr)
    op.build_csr(64)
    rp, cols, vals, diag = op.csr_arrays()
    del vals, diag
    hist = np.zeros(40, dtype=np.int64)
    tiles, wins = (256, 4096), (4096, 20480, 65536, 131072)
    inwin = {(T, W): 0 for T in tiles for W in wins if W > T}
    for r0 in range(0, dim, 1 << 20):
        r1 = min(dim, r0 + (1 << 20))
        rows = np.repeat(np.arange(r0, r1, dtype=np.int64), np.diff(rp[r0:r1 + 1]))
        c = cols[rp[r0]:rp[r1]].astype(np.int64)
        hist += np.bincount(np.floor(np.log2(np.abs(c - rows) + 1)).astype(np.int64), minlength=40)[:40]
        for (T, W) in inwin:
            lo = (rows // T) * T - (W - T) // 2
            inwin[(T, W)] += int(np.count_nonzero((c >= lo) & (c < lo + W)))
        log(f"rows {r1} of {dim}")
    nnz = int(cols.size)
    cum = np.cumsum(hist) / nnz
    print(json.dumps({"workload": f"column distances of the {args.lattice}x{args.lattice} J1-J2 product",
                      "states": int(dim), "off_diagonal_nonzeros": nnz,
                      "share_with_distance_below_2^k": {int(k): round(float(cum[k - 1]), 4)
                                                        for k in range(1, 25)},
                      "share_in_window": {f"tile {T} rows, window {W}": round(v / nnz, 4)
                                          for (T, W), v in inwin.items()}}), flush=True)


def order_model(args):
    """Locality of the 6x6 product's gathers and the L2 misses an LRU model
    of each die's L2 (4 MiB, 16 ways, 128 B lines) predicts for other orders
    of the same matrix: the basis ordered by relabelled sites (rows and
    columns renumbered), reverse Cuthill-McKee, and processing orders of the
    rows (by smallest column, random)."""
    import ctypes as C
    import lsamd
    from lsamd import models
    pl = probes()
    pl.lsamd_probe_l2_model.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_int64, C.c_int, C.c_int,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L = args.lattice
    n, h, S, expr = workload(L)
    b = lsamd.DeviceBasis(n, h, 1, S)
    dim = b.build()
    reps = b.states
    op = lsamd.DeviceOperator(b, expr)
    op.set_csr_packing(False)
    nnz = op.build_csr(64)
    rp, cols, _, _ = op.csr_arrays()
    del op
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    lens = np.diff(rp)
    rows = np.repeat(np.arange(dim, dtype=np.int64), lens)
    dist = np.abs(cols.astype(np.int64) - rows)
    edges = [0, 16, 256, 4096, 65536, 1 << 20, 1 << 24]
    hist = {f"[{a}, {b_})": float(c) / nnz for a, b_, c in zip(edges[:-1], edges[1:], np.histogram(dist, edges)[0])}
    del dist

    def misses(rp_, cols_):
        m, a = C.c_int64(), C.c_int64()
        if pl.lsamd_probe_l2_model(dim, rp_.ctypes.data, cols_.ctypes.data, 8, 4 << 20, 16, 128,
                                   C.byref(m), C.byref(a)) != 0:
            raise SystemExit(pl.lsamd_probe_error().decode())
        return m.value

    def reorder(pi, relabel_columns):
        """Rows in the order pi (new row t is old row pi[t]); columns renumbered too when relabelled."""
        rl = lens[pi]
        rp2 = np.zeros(dim + 1, dtype=np.int64)
        np.cumsum(rl, out=rp2[1:])
        idx = np.repeat(rp[pi] - rp2[:-1], rl) + np.arange(nnz, dtype=np.int64)
        c2 = cols[idx]
        if relabel_columns:
            inv = np.empty(dim, dtype=np.int64)
            inv[pi] = np.arange(dim)
            c2 = inv[c2]
        return rp2, np.ascontiguousarray(c2, dtype=np.uint32)

    def site_order(perm):
        key = np.zeros(dim, dtype=np.uint64)
        for i, p in enumerate(perm):
            key |= ((reps >> np.uint64(p)) & np.uint64(1)) << np.uint64(i)
        return np.argsort(key, kind="stable")

    rc = lambda s_: divmod(s_, L)  # noqa: E731
    orders = {
        "stored (sorted representatives, row-major sites)": None,
        "sites reversed": site_order(list(range(n))[::-1]),
        "sites column-major": site_order([c * L + r for r in range(L) for c in range(L)]),
        "sites snake": site_order([r * L + (c if r % 2 == 0 else L - 1 - c) for r in range(L) for c in range(L)]),
        "sites checkerboard": site_order(sorted(range(n), key=lambda s_: (sum(rc(s_)) % 2, s_))),
    }
    rng = np.random.default_rng(1)
    for q in range(2):
        orders[f"sites random {q}"] = site_order(list(rng.permutation(n)))
    out = {}
    base = misses(rp, cols)
    out["stored (sorted representatives, row-major sites)"] = base
    log(f"stored order: {base / 1e6:.1f} M misses")
    for name, pi in orders.items():
        if pi is None:
            continue
        m = misses(*reorder(pi, True))
        out[name] = m
        log(f"{name}: {m / 1e6:.1f} M misses")
    mincol = np.full(dim, np.iinfo(np.int64).max)
    np.minimum.at(mincol, rows, cols.astype(np.int64))
    out["rows by smallest column (processing order only)"] = misses(*reorder(np.argsort(mincol, kind="stable"), False))
    out["rows random (processing order only)"] = misses(*reorder(rng.permutation(dim), False))
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    A = sp.csr_matrix((np.ones(nnz, dtype=np.int8), cols.astype(np.int32), rp), shape=(dim, dim))
    out["reverse Cuthill-McKee (rows and columns)"] = misses(*reorder(reverse_cuthill_mckee(A, symmetric_mode=True),
                                                                         True))
    for k, v in out.items():
        log(f"{k}: {v / 1e6:.1f} M misses ({v / base:.3f} of the stored order)")
    print(json.dumps({"workload": "order-model", "states": dim, "nonzeros": nnz, "column_distance_fractions": hist,
                      "model": "LRU per die: 8 dies, 4 MiB, 16 ways, 128 B lines; die d takes rows [n d/8, n (d+1)/8)",
                      "x_line_misses": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lattice", type=int, default=6)
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--no-general", action="store_true", help="skip the unpacked and random-coupling products")
    ap.add_argument("--no-reference-abi", action="store_true",
                    help="skip timing ls_chpl_matrix_vector_product with host vectors")
    ap.add_argument("--lanczos-steps", type=int, default=200)
    ap.add_argument("--block-size", type=int, default=2,
                    help="columns of the FP32 block solver for the ground state (a sweep of 1, 2, 4, 8 picked 2)")
    ap.add_argument("--block-cycle", type=int, default=16, help="Krylov blocks per restart cycle")
    ap.add_argument("--block-keep", type=int, default=4, help="Ritz blocks kept at a restart")
    ap.add_argument("--workload", choices=["square6", "chain-sectors", "configs", "copy-calibration",
                                           "gather-ceiling", "order-model", "column-distance"],
                    default="square6")
    ap.add_argument("--copy-mib", type=int, default=2048)
    ap.add_argument("--chain", type=int, default=28)
    ap.add_argument("--sector-blocks", type=int, default=16)
    ap.add_argument("--sector-cycles", type=int, default=40)
    ap.add_argument("--sector-keep", type=int, default=4)
    ap.add_argument("--sector-switch-tol", type=float, default=1e-5,
                    help="relative residual at which the sector eigensolver leaves the FP32 product")
    ap.add_argument("--reference-sectors", default="0+,14+,7-,13+")
    ap.add_argument("--reference-steps", type=int, default=400)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU sharding "
              "is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy-calibration":
        copy_calibration(args)
        return
    if args.workload == "gather-ceiling":
        gather_ceiling(args)
        return
    import lsamd
    from lsamd import _abi
    lib = lsamd.load()
    if lib.ls_amd_device_available() != 0:
        raise SystemExit("no usable gfx950 device: " + _abi.last_error())
    if args.workload == "chain-sectors":
        sectors_bench(args)
        return
    if args.workload == "configs":
        configs_bench(args)
        return
    if args.workload == "order-model":
        order_model(args)
        return
    if args.workload == "column-distance":
        column_distance(args)
        return

    n, h, S, expr = workload(args.lattice)
    basis = lsamd.DeviceBasis(n, h, 1, S)
    lib.ls_amd_synchronize()
    # the first build also allocates the enumeration scratch the later ones
    # reuse; the build rate is the mean of three builds after it
    t0 = time.perf_counter()
    dim = basis.build()
    lib.ls_amd_synchronize()
    t_build_first = time.perf_counter() - t0
    builds = []
    for _ in range(3):
        t0 = time.perf_counter()
        if basis.build() != dim:
            raise SystemExit("basis rebuild changed the number of states")
        lib.ls_amd_synchronize()
        builds.append(time.perf_counter() - t0)
    t_build = float(np.mean(builds))
    candidates = basis.number_candidates
    log(f"basis: {dim} states from {candidates} candidates in {t_build:.4f} s (first build {t_build_first:.4f} s)")

    op = lsamd.DeviceOperator(basis, expr)
    t0 = time.perf_counter()
    nnz_off = op.build_csr(96)
    t_csr = time.perf_counter() - t0
    nnz = nnz_off + dim  # stored off-diagonal entries + the diagonal
    log(f"csr: {nnz_off} off-diagonal nonzeros in {t_csr:.2f} s")
    lanes = op.autotune()
    sched = op.schedule
    log(f"autotune: {sched}")
    nt, die = sched["flags_f64"] & 1, (sched["flags_f64"] >> 1) & 1
    csr_info = op.csr_info
    packed = csr_info["packed"]
    tf = lambda b: "true" if b else "false"  # noqa: E731
    kernel_name = f"csr_f64_kernel<{lanes}, {tf(nt)}, {tf(die)}, {tf(packed)}>"
    if die and sched["flags_f64"] & 4:
        kernel_name += " full grid"
    log(f"csr: value table {'of ' + str(csr_info['table_size']) + ' values' if packed else 'not used'}")

    rng = np.random.default_rng(42)
    x = lsamd.DeviceArray.from_numpy(rng.standard_normal(dim))
    y = lsamd.DeviceArray(8 * dim)

    # ---- timed FP64 steps ----
    for _ in range(args.warmup):
        op.csr_matvec(x, y)
    lib.ls_amd_synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.csr_matvec(x, y)
        kernel_ms.append(lib.ls_amd_last_kernel_ms(0))
    lib.ls_amd_synchronize()
    t_steps = time.perf_counter() - t0
    ms_per_step = 1e3 * t_steps / args.steps
    value = nnz * args.steps / t_steps
    k_ms = float(np.mean(kernel_ms))
    bytes_per_launch = 20.0 * nnz_off + 32.0 * dim
    achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9
    traffic, traffic_src = profiled_traffic(f"square{args.lattice}", kernel_name)

    # ---- FP32 block of 8 ----
    X8 = lsamd.DeviceArray.from_numpy(rng.standard_normal((dim, 8)).astype(np.float32))
    Y8 = lsamd.DeviceArray(4 * 8 * dim)
    for _ in range(args.warmup):
        op.csr_matvec_f32(8, X8, Y8)
    lib.ls_amd_synchronize()
    b_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.csr_matvec_f32(8, X8, Y8)
        b_ms.append(lib.ls_amd_last_kernel_ms(1))
    lib.ls_amd_synchronize()
    t_b = time.perf_counter() - t0
    block_rate = 8.0 * nnz * args.steps / t_b
    b_ms_mean = float(np.mean(b_ms))
    # compulsory bytes: the stored CSR stream (4 B per nonzero with the value
    # table, else 8), row pointer 8 and diagonal 4 per row, x read once and y
    # written once (8 FP32 columns: 32 B each per row)
    block_bytes = (4.0 if packed else 8.0) * nnz_off + dim * (8 + 4 + 64)
    block_gbs = block_bytes / (b_ms_mean * 1e-3) / 1e9
    block_gather_gbs = (40.0 * nnz_off + dim * (8 + 4 + 64)) / (b_ms_mean * 1e-3) / 1e9

    # ---- the reference-named entry point: host x and y, copied over PCIe on every call ----
    ref_abi = None
    if not args.no_reference_abi:
        sb = lsamd.SpinBasis(n, h, 1, S)
        sb.build()
        hop = lsamd.Operator(sb, expr)
        xh = rng.standard_normal(dim)
        yh = np.zeros(dim)
        t0 = time.perf_counter()
        hop.matvec(xh, out=yh)  # the first call uploads the basis and builds the cached CSR
        t_first = time.perf_counter() - t0
        calls = 5
        t0 = time.perf_counter()
        for _ in range(calls):
            hop.matvec(xh, out=yh)
        t_call = (time.perf_counter() - t0) / calls
        dxh = lsamd.DeviceArray.from_numpy(xh)
        op.csr_matvec(dxh, y)
        yd = y.to_numpy(np.float64, (dim,))
        ref_abi = {"entry_point": "ls_chpl_matrix_vector_product", "seconds_per_call": t_call,
                   "nonzeros_per_s_pcie_inclusive": nnz / t_call, "first_call_seconds": t_first,
                   "max_rel_diff_to_device_product": float(np.abs(yh - yd).max() / np.abs(yd).max()),
                   "note": "x and y are host arrays, copied to and from the GPU on every call"}
        log(f"reference entry point: {1e3 * t_call:.1f} ms per call with host x and y "
            f"({nnz / t_call:.3e} nonzeros/s), first call {t_first:.1f} s")
        del hop, sb, dxh

    # ---- the general case: no value table ----
    general = None if args.no_general else general_case(lib, basis, args.lattice, dim, args.steps, rng)

    # ---- on-the-fly fused FP64 (no CSR) ----
    t0 = time.perf_counter()
    op.matvec_otf(x, y)
    lib.ls_amd_synchronize()
    t_otf = time.perf_counter() - t0

    # ---- HDF5 interchange of the basis and a vector ----
    h5 = hdf5_io(basis, x, dim)

    # ---- ground state: plain FP64 vs the FP32 block solver + FP64 final cycles ----
    e0 = {}
    if not args.no_lanczos:
        r64 = op.lanczos("fp64", max_steps=args.lanczos_steps, k=1, tol=1e-12)
        rb = op.block_eigensolve("fp32", block_size=args.block_size, nev=1, max_blocks=args.block_cycle,
                                 keep_blocks=args.block_keep, max_cycles=40, tol=1e-8, switch_tol=1e-5)
        e_b = float(rb["eigenvalues"][0])
        d = abs(e_b - r64["eigenvalues"][0]) / abs(r64["eigenvalues"][0])
        e0 = {"E0_fp64": r64["eigenvalues"][0], "fp64_steps": r64["steps"],
              "fp64_residual": r64["residual_norm"], "fp64_seconds": r64["total_seconds"],
              "E0_fp32_block_then_fp64": e_b, "block_size": args.block_size, "block_cycle": args.block_cycle,
              "block_keep": args.block_keep, "block_cycles": rb["cycles"],
              "block_fp64_cycles": rb["fp64_cycles"], "block_products": rb["products"],
              "block_relative_residual": float(rb["residuals"][0]), "block_seconds": rb["total_seconds"],
              "block_product_seconds": rb["product_seconds"],
              "speedup_vs_fp64": r64["total_seconds"] / rb["total_seconds"], "dE0_rel": d}
        log(f"E0 fp64 {r64['eigenvalues'][0]:.15f} ({r64['steps']} steps, {r64['total_seconds']:.2f} s), "
            f"fp32 block of {args.block_size} -> fp64 {e_b:.15f} ({rb['cycles']} cycles, {rb['fp64_cycles']} fp64, "
            f"{rb['products']} products, {rb['total_seconds']:.2f} s), |dE0/E0| = {d:.2e}")
        # the same solve on the FP32 block-of-8 product that fp32_block8 times
        r8 = op.block_eigensolve("fp32", block_size=8, nev=1, max_blocks=args.block_cycle,
                                 keep_blocks=args.block_keep, max_cycles=40, tol=1e-8, switch_tol=1e-5)
        e_8 = float(r8["eigenvalues"][0])
        d8 = abs(e_8 - r64["eigenvalues"][0]) / abs(r64["eigenvalues"][0])
        e0["block8"] = {"E0_fp32_block_then_fp64": e_8, "dE0_rel": d8, "block_size": 8,
                        "block_cycle": args.block_cycle, "block_keep": args.block_keep, "cycles": r8["cycles"],
                        "fp64_cycles": r8["fp64_cycles"], "products": r8["products"],
                        "relative_residual": float(r8["residuals"][0]), "seconds": r8["total_seconds"],
                        "product_seconds": r8["product_seconds"]}
        log(f"fp32 block of 8 -> fp64 {e_8:.15f} ({r8['cycles']} cycles, {r8['fp64_cycles']} fp64, "
            f"{r8['products']} products, {r8['total_seconds']:.2f} s), |dE0/E0| = {d8:.2e}")

    line = {
        "metric": "H·x nonzeros/s + basis states/s, 6×6 J1-J2; FP32-block vs FP64 at ΔE0≤1e-10",
        "value": value,
        "unit": "nonzeros/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded N(0,1) x (seed 42); Hamiltonian and symmetries of "
                "heisenberg_square_6x6.yaml plus J2 = 0.5",
        "config": {"workload": f"J1-J2 {args.lattice}x{args.lattice} square lattice, J2=0.5, Hamming weight {h}, "
                               "T x C4v x Z2 (576 elements), FP64 CSR H x",
                   "states": dim, "nonzeros": nnz, "off_diagonal_nonzeros": nnz_off,
                   "schedule": sched, "value_table": csr_info["table_size"] if packed else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     **traffic_rate(traffic, k_ms),
                     "kernel": kernel_name, "kernel_ms": k_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "basis_build": {"states": dim, "candidates": candidates, "seconds": t_build,
                        "states_per_s": dim / t_build, "candidates_per_s": candidates / t_build,
                        "first_build_seconds": t_build_first, "timed": "mean of 3 builds after the first"},
        "csr_build_seconds": t_csr,
        "reference_abi_host_vectors": ref_abi,
        "fp32_block8": {"nonzero_rhs_per_s": block_rate, "speedup_vs_fp64": block_rate / value,
                        "kernel_ms": b_ms_mean, "achieved_GBs": block_gbs, "frac": block_gbs / HBM_PEAK_GBS,
                        "bytes_model": "compulsory: CSR stream once, x read once, y written once",
                        "gather_model_GBs": block_gather_gbs,
                        "gather_model_frac": block_gather_gbs / HBM_PEAK_GBS,
                        "gather_model": "per stored nonzero 8 B (column, value) + one 32 B row of x; "
                                        "per row 8 + 4 + 64 B"},
        "general_case": general,
        "otf_fp64": {"seconds": t_otf, "nonzeros_per_s": nnz / t_otf},
        "hdf5_io": h5,
        "ground_state": e0,
    }
    if not args.no_cpu_baseline:
        states = basis.states
        line["cpu_baseline"] = cpu_baseline(S, n, h, states, expr, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

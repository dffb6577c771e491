#!/usr/bin/env python3
"""Groth16 prove benchmark on one MI355X (gfx950).

Workload (N=1): one plain Groth16 proof of the 100x100 matmul R1CS with 10%
non-zero A/B entries (uniform Fr), 1,010,000 live constraints in a 2^20
domain, 2^20 columns, on BN254.  A step is Backend::prove (witness upload,
SpMV x3, the quotient chain, four G1 MSMs and one G2 MSM) followed by
blind_proof with r, s from the seed -- the reference's groth16.prove.total
scope (protocols/groth16/app/prove.cpp:79-88), whose Backend::prove receives
the witness as host memory and uploads it (backends/hip/src/backend.hip:
302-307).  The circuit and proving key are resident in HBM before the timed
region (prepare is outside that scope); the keys come from the GPU setup with
seeded trapdoors, and the timed proof is verified.  The same step with the
witness already in HBM is reported beside it.

Two more legs (skippable): `batch` proves a heterogeneous mix of 64 proofs
(16 circuit shapes, domains 2^9 .. 2^18) plainly and as one batch, and
`serving` replays a Poisson request stream (the mix plus 10% 2^20 requests)
through the plain one-at-a-time path and through continuous batching,
reporting p50/p99 latency and proofs/s.

Prints one JSON line.  `value` is proofs/s of the whole job; `ms_per_step`
is the prove latency.  The CPU baseline times the repository's C restatement
of the reference CPU backend on this host, at 1 thread and on all host cores
(see `cpu_baseline`).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nkr-zkp_amd"))

import numpy as np  # noqa: E402

import circuits  # noqa: E402
import autotune  # noqa: E402
import monitor  # noqa: E402
import nkr_g16 as N  # noqa: E402

SEED = 0x4E4B52
HBM_PEAK_GBS = 8000.0                   # MI355X HBM3E, published
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9   # 256 CUs x 64 lanes x 2.4 GHz; v_mad_u64_u32 issues at full rate
MADS_PER_FQ_MUL = 136                   # 2*8^2 + 8 for an 8x32-bit Montgomery product
FQ_MUL_PEAK = VALU_PEAK_LANE_OPS / MADS_PER_FQ_MUL
ONE_MONT = np.array([((1 << 256) % circuits.R >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for k in range(4)],
                    dtype=np.uint64)    # 1 in Montgomery form (R = 2^256 mod r)
STAGES = ["spmv", "ntt", "quotient", "msm_sort", "msm_accumulate_g1", "msm_fold", "finish_host", "msm_accumulate_g2"]


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pmc_workload(size: int, density: str) -> int:
    """Workload for `rocprofv3 --pmc` passes (one counter per run): three
    device-to-device copies of 256 MiB (k_copy16 reads and writes exactly
    that much, which calibrates FETCH_SIZE / WRITE_SIZE), then two proves of
    the benchmark circuit with the witness in HBM."""
    L = N.lib()
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    N._check(L.nkr_dev_alloc(nbytes, ctypes.byref(a)))
    N._check(L.nkr_dev_alloc(nbytes, ctypes.byref(b)))
    for _ in range(3):
        N._check(L.nkr_dev_copy(b, a, nbytes, None))
    N._check(L.nkr_device_synchronize())
    log("pmc workload: copies done")
    circ = circuits.matmul(size, SEED, density=density, wide=True)
    pk = N.setup(circ, circuits.random_scalars(5, SEED ^ 0x5E7))
    pv = N.Prover(circ, pk)
    wit = np.ascontiguousarray(circ.witness)
    d = ctypes.c_void_p()
    N._check(L.nkr_dev_alloc(wit.nbytes, ctypes.byref(d)))
    N._check(L.nkr_memcpy_htod(d, N._p(wit), wit.nbytes))
    rs = np.ascontiguousarray(circuits.random_scalars(2, SEED ^ 0x72))
    raw = np.zeros(72, dtype=np.uint64)
    proof = np.zeros(48, dtype=np.uint64)
    for _ in range(2):
        N._check(L.nkr_g16_prove_blind_dev(pv.h, d, N._p(rs[0]), N._p(rs[1]), N._p(raw), N._p(proof)))
    N._check(L.nkr_device_synchronize())
    log("pmc workload: proves done")
    for p in (a, b, d):
        N._check(L.nkr_dev_free(p))
    pv.close()
    return 0


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_prove_matmul100_sparse_2p20.md")


def measured_traffic(kernel: str):
    """HBM bytes of one accumulation launch ("g1": k_accumulate<Fq>, "g2":
    k_accumulate<Fq2>) of this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary scaled by the copy-kernel calibration;
    None if the summary is absent."""
    key = f"traffic_bytes_{kernel}_accumulate:"
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(key):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def percentile(xs, q: float) -> float:
    s = sorted(xs)
    return s[min(len(s) - 1, int(round(q / 100.0 * (len(s) - 1))))]


def host_cpu() -> dict:
    """The host CPU this run's baseline ran on: model name, logical CPUs of the
    machine and CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(circ, pk, witness_np, h_host) -> dict:
    """The oracle (C restatement of the reference CPU backend: CSR SpMV, serial
    radix-2 FFT, libff BDLO12 Pippenger with complete Jacobian additions) on
    this host, in two legs:
      1 thread (the reference's policy, OMP_NUM_THREADS=1): SpMV x3, one 2^20
      iFFT (x7: the chain's 3 iFFT, 3 cosetFFT and 1 icosetFFT cost the same),
      one witness G1 MSM (x3: ar, bs1 and zk share the scalars), the G2 MSM
      and the H MSM, each at full size;
      all host cores: the whole proof (or_prove_mt, MSMs chunked over threads
      as variable_base_msm does), timed end to end."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    n = circ.rows
    t = {}
    t0 = time.perf_counter()
    for k in ("a", "b", "c"):
        O.spmv(circ.rows, circ.cols, *getattr(circ, k), witness_np)
    t["spmv_x3"] = time.perf_counter() - t0
    a = witness_np[:n].copy()
    t0 = time.perf_counter()
    O.fft(a, O.IFFT)
    t["ntt_x7"] = 7 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    O.msm_g1(pk.a_query, witness_np)
    t["msm_g1_witness_x3"] = 3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    O.msm_g2(pk.b_g2_query, witness_np)
    t["msm_g2_witness"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.msm_g1(pk.h_query, h_host)
    t["msm_g1_h"] = time.perf_counter() - t0
    total = sum(t.values())
    log(f"CPU baseline, 1 thread: {total:.1f} s per proof")
    threads = O.host_threads()
    t0 = time.perf_counter()
    O.prove_mt(circ, pk, witness_np, threads=threads)
    all_s = time.perf_counter() - t0
    log(f"CPU baseline, {threads} threads: {all_s:.2f} s per proof")
    return {
        "value": round(1.0 / total, 6), "unit": "proofs/s", "cores": 1, "kind": "port",
        "prove_s": round(total, 3),
        "stage_s": {k: round(v, 3) for k, v in t.items()},
        "sample": ("oracle (C restatement of the reference CPU backend: CSR SpMV, serial radix-2 FFT, libff "
                   "BDLO12), 1 thread, this proof at full size: SpMV x3, one 2^20 iFFT (x7), one witness G1 MSM "
                   "(x3), the G2 MSM and the H MSM"),
        "all_cores": {"value": round(1.0 / all_s, 6), "unit": "proofs/s", "cores": threads, "prove_s": round(all_s, 3),
                      "sample": "the whole proof once (or_prove_mt: MSMs chunked over the threads)"},
        "host": host_cpu(),
    }


def operators_leg(reps: int) -> list:
    """Standalone operators at the sizes of the reference's operator
    configuration, inputs resident in HBM: forward Fr NTT at 2^16, 2^20 and
    2^24 (nkr_fr_ntt_dev, in place); CSR SpMV over 2^20 rows with 3 non-zeros
    per row on average (nkr_spmv_run_dev, matrix uploaded beforehand); G1 MSM at 2^16, 2^20 and 2^24 with uniform
    and witness-like (half 0/1) scalars and G2 MSM at 2^18 with uniform
    scalars (nkr_msm_run_dev, window-shifted bases prepared beforehand, the
    host's final combination included).  Bases are distinct multiples of
    points from the GPU setup of a small circuit.  Time per call = wall time
    of `reps` back-to-back calls between device synchronisations."""
    lib = N.lib()
    rows = []

    def to_dev(a):
        ptr = ctypes.c_void_p()
        N._check(lib.nkr_dev_alloc(a.nbytes, ctypes.byref(ptr)))
        N._check(lib.nkr_memcpy_htod(ptr, N._p(a), a.nbytes))
        return ptr

    def timed(fn):
        fn()
        N._check(lib.nkr_device_synchronize())
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        N._check(lib.nkr_device_synchronize())
        return 1e3 * (time.perf_counter() - t0) / reps

    for log_n in (16, 20, 24):
        n = 1 << log_n
        buf = to_dev(circuits.random_scalars(n, SEED + log_n))
        ms = timed(lambda: N._check(lib.nkr_fr_ntt_dev(buf, log_n, N.FFT, None)))
        N._check(lib.nkr_dev_free(buf))
        products = n // 2 * log_n
        rows.append({"op": "fr_ntt_forward", "log_n": log_n, "ms": round(ms, 4),
                     "m_elements_per_s": round(n / ms / 1e3, 1),
                     "g_products_per_s": round(products / ms / 1e6, 2),
                     "valu_frac": round(products / ms * 1e3 / FQ_MUL_PEAK, 4)})
        log(f"operators: NTT 2^{log_n} {ms:.3f} ms")

    rows_n = 1 << 20
    rng = np.random.default_rng(SEED)
    per_row = rng.poisson(3, rows_n)
    row_ptr = np.concatenate([[0], np.cumsum(per_row)]).astype(np.uint32)
    nnz = int(row_ptr[-1])
    sp = N.Spmv(rows_n, rows_n, row_ptr, rng.integers(0, rows_n, nnz).astype(np.uint32),
                circuits.random_scalars(nnz, SEED ^ 0x5B))
    z = to_dev(circuits.random_scalars(rows_n, SEED ^ 0x5C))
    y = to_dev(np.zeros((rows_n, 4), dtype=np.uint64))
    ms = timed(lambda: sp.run_dev(z.value, y.value))
    sp.close()
    N._check(lib.nkr_dev_free(z))
    N._check(lib.nkr_dev_free(y))
    alg = 36 * nnz + 36 * rows_n + 32 * rows_n  # values + columns, offsets + output, gathered z (once)
    rows.append({"op": "spmv", "rows": rows_n, "nnz": nnz, "ms": round(ms, 4),
                 "alg_gb_per_s": round(alg / ms / 1e6, 1), "hbm_frac": round(alg / ms / 1e6 / HBM_PEAK_GBS, 4)})
    log(f"operators: SpMV 2^20 rows, {nnz} non-zeros {ms:.3f} ms")

    small = circuits.matmul(2, SEED, wide=True)
    keys = N.setup(small, circuits.random_scalars(5, SEED ^ 0x0B5))
    for g2, sizes in ((False, (16, 20, 24)), (True, (18,))):
        base = keys.beta_g2 if g2 else keys.alpha_g1
        for log_n in sizes:
            n = 1 << log_n
            bases = N.fixed_base_mul(base, circuits.random_scalars(n, SEED ^ (0xBA5E + log_n)), g2=g2)
            m = N.Msm(bases, g2=g2)
            del bases
            s = circuits.random_scalars(n, SEED ^ (0x5CA1 + log_n))
            variants = ["uniform"] if g2 else ["uniform", "witness"]
            for variant in variants:
                if variant == "witness":
                    rng = np.random.default_rng(SEED + log_n)
                    pick = rng.random(n) < 0.5
                    s[pick] = 0
                    s[pick & (rng.random(n) < 0.5)] = ONE_MONT
                buf = to_dev(s)
                ms = timed(lambda: m.run_dev(buf.value))
                N._check(lib.nkr_dev_free(buf))
                rows.append({"op": "g2_msm" if g2 else "g1_msm", "log_n": log_n, "scalars": variant,
                             "window_bits": m.window_bits, "ms": round(ms, 4),
                             "m_terms_per_s": round(n / ms / 1e3, 1)})
                log(f"operators: {'G2' if g2 else 'G1'} MSM 2^{log_n} {variant} {ms:.3f} ms")
            m.close()
    return rows


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=100, help="matmul dimension s (100 -> 2^20 domain)")
    ap.add_argument("--density", default="sparse")
    ap.add_argument("--circuit", choices=["matmul", "poseidon"], default="matmul",
                    help="matmul (the default workload) or the Poseidon-style batch: --hashes independent "
                         "width-3 hashes, folded rounds (~240 constraints each; 4096 -> 2^20 domain)")
    ap.add_argument("--hashes", type=int, default=4096)
    ap.add_argument("--tuned", action="store_true",
                    help="build the context with the shipped MI355X tuning table's entry for this domain "
                         "size and witness density (autotune.tuning_for); the default uses library defaults")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain mix")
    ap.add_argument("--batch-per-shape", type=int, default=4, help="witnesses per circuit shape in the mix")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--qps", default="200,2000",
                    help="Poisson arrival rates of the serving leg (comma list): one below the plain path's "
                         "capacity for latency, one above both paths' capacity for throughput")
    ap.add_argument("--requests", type=int, default=300, help="requests in the serving leg")
    ap.add_argument("--verify-all", action="store_true", help="also verify the served 2^20 proofs")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="only run the counter-collection workload: three 256 MiB device copies (the "
                         "calibration kernel), then two proves of the 2^20 circuit")
    ap.add_argument("--operators", action="store_true",
                    help="only time the standalone operators (NTT, G1/G2 MSM) at the operator-configuration "
                         "sizes and print them as one JSON line")
    ap.add_argument("--op-reps", type=int, default=10)
    ap.add_argument("--no-util", action="store_true",
                    help="skip the utilisation qualification of the 2^20 proof (concurrency 1..16)")
    ap.add_argument("--util-window", type=float, default=1.5,
                    help="seconds of back-to-back calls per concurrency in the qualification")
    ap.add_argument("--util-max-p", type=int, default=16)
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the batch leg (the 64-proof mix, plain and batched) and print it as one JSON "
                         "line: the workload of the batch kernel trace")
    args = ap.parse_args()
    if args.pmc_workload:
        return pmc_workload(args.size, args.density)
    if args.batch_only:
        shapes, items = build_mix(args.batch_per_shape)
        print(json.dumps({"batch": batch_leg(shapes, items, args.batch_per_shape, args.batch_reps),
                          "data": "synthetic (seeded)"}), flush=True)
        for pv in list({id(it[2]): it[2] for it in items}.values()):
            pv.close()
        return 0
    if args.operators:
        print(json.dumps({"operators": operators_leg(args.op_reps), "data": "synthetic (seeded)",
                          "inputs": "resident in HBM"}), flush=True)
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not part of "
                                   "this build (proofs shard across GPUs as independent work lists)"}))
        return 2

    if args.circuit == "poseidon":
        log(f"generating {args.hashes} folded Poseidon-style hashes seed={SEED:#x}")
        circ = circuits.generate("poseidon", args.hashes, SEED, folded=True)
        workload = f"groth16_prove_poseidon{args.hashes}_folded"
    else:
        log(f"generating matmul {args.size} ({args.density}) seed={SEED:#x}")
        circ = circuits.matmul(args.size, SEED, density=args.density, wide=True)
        workload = f"groth16_prove_matmul{args.size}_{args.density}"
    log(f"rows={circ.rows} cols={circ.cols} nnz=({circ.a[1].size},{circ.b[1].size},{circ.c[1].size})")
    t0 = time.perf_counter()
    pk = N.setup(circ, circuits.random_scalars(5, SEED ^ 0x5E7))
    log(f"proving and verification keys generated on the GPU ({time.perf_counter() - t0:.2f}s)")
    t0 = time.perf_counter()
    tuning = autotune.tuning_for(circ, circ.witness) if args.tuned else None
    if args.tuned:
        log(f"tuning from the shipped table: {tuning or 'library defaults'}")
    prover = N.Prover(circ, pk, tuning=tuning)
    prep_s = time.perf_counter() - t0
    log(f"prepare {prep_s:.2f}s, device bytes {prover.device_bytes() / 2**30:.2f} GiB")

    wit = np.ascontiguousarray(circ.witness)
    raw = np.zeros(72, dtype=np.uint64)
    proof = np.zeros(48, dtype=np.uint64)
    rs = np.ascontiguousarray(circuits.random_scalars(2, SEED ^ 0x72))
    r_p, s_p = rs[0].ctypes.data, rs[1].ctypes.data
    dptr = ctypes.c_void_p()
    N._check(N.lib().nkr_dev_alloc(wit.nbytes, ctypes.byref(dptr)))
    N._check(N.lib().nkr_memcpy_htod(dptr, N._p(wit), wit.nbytes))

    def step():
        # Backend::prove + blind_proof (groth16.prove.total): the witness is
        # ordinary host memory, as Backend::prove receives it, and its upload
        # is part of the step
        N._check(N.lib().nkr_g16_prove_blind(prover.h, N._p(wit), r_p, s_p, N._p(raw), N._p(proof)))

    def step_dev():
        # the same with the witness already resident in HBM
        N._check(N.lib().nkr_g16_prove_blind_dev(prover.h, dptr, r_p, s_p, N._p(raw), N._p(proof)))

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        N._check(N.lib().nkr_device_synchronize())
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        N._check(N.lib().nkr_device_synchronize())
        return 1000.0 * (time.perf_counter() - t0) / args.steps

    ms_per_step = timed(step)
    first_raw, first_proof = raw.copy(), proof.copy()
    ms_dev = timed(step_dev)
    deterministic = bool(np.array_equal(first_raw, raw) and np.array_equal(first_proof, proof))
    # the fused blinding equals blind_proof applied to the raw proof
    blind_consistent = bool(np.array_equal(prover.blind(raw, rs[0], rs[1]), proof))
    N._check(N.lib().nkr_dev_free(dptr))
    # bytes the packed upload sends: values other than 0 and 1, masks and offsets
    trivial = ~wit.any(axis=1) | (wit == ONE_MONT).all(axis=1)
    upload_bytes = int((~trivial).sum()) * 32 + ((circ.cols + 63) // 64) * 20
    log(f"{ms_per_step:.3f} ms/prove with the host witness, {ms_dev:.3f} ms with it in HBM")

    # Separate instrumented pass: per-stage device time (HIP events on the
    # streams the kernels run on; the two streams overlap) and entry counts.
    N._check(N.lib().nkr_g16_profile(prover.h, 1))
    for _ in range(args.steps):
        step()
    N._check(N.lib().nkr_device_synchronize())
    stage_ms = (ctypes.c_double * len(STAGES))()
    counters = (ctypes.c_uint64 * 4)()
    N._check(N.lib().nkr_g16_profile_read(prover.h, stage_ms, counters))
    N._check(N.lib().nkr_g16_profile(prover.h, 0))
    stages = {k: round(stage_ms[i] / args.steps, 3) for i, k in enumerate(STAGES)}
    log(f"stages (device time per prove, overlapping streams) {stages}")

    # blinding alone (host, stateless entry point)
    t0 = time.perf_counter()
    prover.blind(raw, rs[0], rs[1])
    blind_ms = 1000.0 * (time.perf_counter() - t0)

    # Integer-VALU roof measured on this device: v_mad_u64_u32 issue rate,
    # and the rate of this library's Montgomery product.
    mad_s, mul_s = ctypes.c_double(), ctypes.c_double()
    N._check(N.lib().nkr_valu_probe(0, ctypes.byref(mad_s), ctypes.byref(mul_s)))
    mul_peak = mad_s.value / MADS_PER_FQ_MUL
    log(f"VALU probe: {mad_s.value / 1e12:.2f} T v_mad_u64_u32/s ({mul_peak / 1e9:.1f} G products/s at 136 MADs "
        f"each), library product {mul_s.value / 1e9:.1f} G/s")

    # Dominant kernel: the G1 MSM bucket accumulation (one k_accumulate<Fq>
    # launch per prove for ar, bs1, zk and H).  Per entry (one non-zero
    # signed digit of one term) it reads a 4-byte sorted index and one affine
    # base (64 B) and performs one mixed XYZZ addition (8M + 2S = 10 Fq
    # products).  The G2 accumulation (bs2: 128 B bases, 30 Fq products per
    # entry with 3-product Fq2 Karatsuba) is reported beside it.
    proves = max(1, counters[0])
    e_w = counters[2] / proves          # witness schedule: 3 G1 MSMs + 1 G2 MSM
    e_h = counters[3] / proves          # H schedule: 1 G1 MSM
    g1_ms = stage_ms[4] / args.steps
    g2_ms = stage_ms[7] / args.steps
    kern = {}
    for name, ms, entries, per_bytes, per_mul in (("g1", g1_ms, 3 * e_w + e_h, 68, 10), ("g2", g2_ms, e_w, 132, 30)):
        muls = entries * per_mul
        kern[name] = {"launch_ms": round(ms, 4), "entries": int(entries), "alg_fq_mul": int(muls),
                      "alg_bytes": int(entries * per_bytes),
                      "g_fq_mul_per_s": round(muls / (ms * 1e-3) / 1e9, 2),
                      "hbm_gb_per_s": round(entries * per_bytes / (ms * 1e-3) / 1e9, 1)}
    g1 = kern["g1"]
    traffic = measured_traffic("g1") if (args.circuit, args.size, args.density) == ("matmul", 100, "sparse") \
        else None
    # witness windows stop at 15 bits; a tuned context may set either width
    cw = (tuning or {}).get("witness_window_bits") or min(15, _msm_window(circ.cols))
    ch = (tuning or {}).get("h_window_bits") or _msm_window(circ.rows)
    result = {
        "metric": "Groth16 prove latency @2^20 constraints; batched proofs/s vs plain (bit-exact)",
        "value": round(1000.0 / ms_per_step, 4),
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 limbs (BN254 Fr/Fq Montgomery, exact)",
        "data": f"synthetic (seeded {args.circuit} R1CS and witness; keys from the GPU setup with seeded trapdoors)",
        "config": {"workload": f"{workload}_domain_2^{circ.rows.bit_length() - 1}",
                   "rows": circ.rows, "cols": circ.cols, "path": "plain (one proof per call)",
                   "step": "nkr_g16_prove_blind: host witness uploaded inside the step, prove, blind_proof",
                   "tuning": tuning or {},
                   "msm_window_bits": {"witness": cw, "h": ch}},
        "ms_per_step_witness_in_hbm": round(ms_dev, 3),
        "witness_upload": {"witness_bytes": int(wit.nbytes), "pcie_bytes": upload_bytes,
                           "ms_over_hbm_resident": round(ms_per_step - ms_dev, 3)},
        "stages_ms": stages,
        "blind_ms_host_stateless": round(blind_ms, 3),
        "deterministic": deterministic,
        "fused_blind_matches_blind_proof": blind_consistent,
        "roofline": {
            "kernel": "k_accumulate<Fq> (G1 MSM bucket accumulation: ar, bs1, zk and H in one launch)",
            "bound": "valu",
            "achieved": g1["g_fq_mul_per_s"], "peak": round(mul_peak / 1e9, 1), "unit": "G Fq-mul/s",
            "frac": round(g1["g_fq_mul_per_s"] * 1e9 / mul_peak, 4),
            "traffic": traffic,
            "peak_source": (f"nkr_valu_probe in this run: {mad_s.value / 1e12:.2f} T v_mad_u64_u32/s over the "
                            f"device / {MADS_PER_FQ_MUL} MADs per Montgomery product"),
            "library_product_rate": round(mul_s.value / 1e9, 1),
            "frac_of_library_product_rate": round(g1["g_fq_mul_per_s"] * 1e9 / mul_s.value, 4),
            "launch_ms": g1["launch_ms"], "alg_fq_mul": g1["alg_fq_mul"],
            "alg_bytes": g1["alg_bytes"], "hbm_gb_per_s": g1["hbm_gb_per_s"],
            "hbm_frac": round(g1["hbm_gb_per_s"] / HBM_PEAK_GBS, 4),
            "g2": kern["g2"],
        },
    }
    result["verified"] = bool(N.verify(pk, wit[1:1 + circ.public_inputs], proof))
    if not args.no_util:
        result["qualification"] = qualification_leg(prover, wit, rs, proof.copy(), args.util_window,
                                                    args.util_max_p)
    h = None
    if not args.no_cpu_baseline:
        _, (_, _, _, h) = prover.prove(wit, intermediates=True)
    mix_provers = []
    if not (args.no_batch and args.no_serve):
        shapes, items = build_mix(args.batch_per_shape)
        mix_provers = list({id(it[2]): it[2] for it in items}.values())
        if not args.no_batch:
            result["batch"] = batch_leg(shapes, items, args.batch_per_shape, args.batch_reps)
        if not args.no_serve:
            result["serving"] = [serving_leg(items, (prover, pk, circ, wit), float(q), args.requests,
                                             args.verify_all) for q in args.qps.split(",")]
    for pv in mix_provers:
        pv.close()
    prover.close()
    if not args.no_cpu_baseline:
        log("timing the CPU baseline sample")
        result["cpu_baseline"] = cpu_baseline(circ, pk, wit, h)
    print(json.dumps(result), flush=True)
    return 0


def _msm_window(n: int) -> int:
    """MsmPlan::choose: ~3/4 log2 n in [4, 15] (16 from 2^20 terms up), moved
    to a neighbour when the top window would keep fewer than 4 of the
    scalar's 254 bits."""
    lg = max(0, (n - 1).bit_length())
    c = 16 if lg >= 20 else max(4, min(15, (lg * 3 + 2) // 4))
    top = lambda cc: 254 - cc * ((255 + cc - 1) // cc - 1)
    if top(c) < 4:
        if c < 16 and top(c + 1) >= 4:
            c += 1
        elif c > 4 and top(c - 1) >= 4:
            c -= 1
    return c


def build_mix(per_shape: int):
    """The heterogeneous mix as prepared contexts: [(name, circ, prover, keys, witness, r, s)]."""
    log(f"batch mix: generating 16 circuit shapes x {per_shape} witnesses")
    t0 = time.perf_counter()
    mix = circuits.batch_mix(SEED, per_shape)
    gen_s = time.perf_counter() - t0
    items = []
    for k, (name, circ, wits) in enumerate(mix):
        keys = N.setup(circ, circuits.random_scalars(5, SEED + 17 * k))
        pv = N.Prover(circ, keys)
        for j, w in enumerate(wits):
            rs = circuits.random_scalars(2, SEED ^ (k << 8) ^ j)
            items.append((name, circ, pv, keys, np.ascontiguousarray(w), rs[0].copy(), rs[1].copy()))
    log(f"batch mix: {len(items)} proofs ready ({gen_s:.1f}s to generate), rows 2^"
        f"{min(c.rows for _, c, *_ in items).bit_length() - 1} .. 2^{max(c.rows for _, c, *_ in items).bit_length() - 1}")
    return [name for name, _, _ in mix], items


def batch_leg(shapes, items, per_shape: int, reps: int) -> dict:
    """Batched vs plain proofs/s over the heterogeneous mix (16 circuit shapes,
    domains 2^9 .. 2^18, per_shape witnesses each).  Plain: one
    nkr_g16_prove_blind per proof, back to back.  Batched: the same proofs in
    one nkr_g16_batch_prove call, the repeated composition replayed as a
    captured HIP graph (the eager launch sequence timed beside it).  Both
    time host-resident witnesses in and
    blinded proofs out; every batched proof must equal its plain proof bit for
    bit and verify."""
    plain_out = [pv.prove_blind(w, r, s) for (_, _, pv, _, w, r, s) in items]  # warm-up + reference results
    t0 = time.perf_counter()
    for _ in range(reps):
        for (_, _, pv, _, w, r, s) in items:
            pv.prove_blind(w, r, s)
    plain_s = (time.perf_counter() - t0) / reps

    batch = N.Batch()
    args = [(pv, w, r, s) for (_, _, pv, _, w, r, s) in items]
    # eager launch sequence (graph capture off)
    batch.set_graphs(False)
    outs_e = batch.prove(args)  # warm-up (work buffers grow here)
    t0 = time.perf_counter()
    for _ in range(reps):
        outs_e = batch.prove(args)
    eager_s = (time.perf_counter() - t0) / reps
    # the repeated composition as a captured HIP graph: one eager call records
    # the composition, the next captures and replays, the timed calls replay
    batch.set_graphs(True)
    batch.prove(args)
    batch.prove(args)
    graph = batch.graph_active()
    t0 = time.perf_counter()
    for _ in range(reps):
        outs = batch.prove(args)
    batch_s = (time.perf_counter() - t0) / reps
    exact = all(np.array_equal(o[0], p[0]) and np.array_equal(o[1], p[1]) for o, p in zip(outs, plain_out))
    exact = exact and all(np.array_equal(o[0], p[0]) and np.array_equal(o[1], p[1])
                          for o, p in zip(outs_e, plain_out))
    verified = all(N.verify(keys, w[1:1 + circ.public_inputs], o[1])
                   for (_, circ, _, keys, w, _, _), o in zip(items, outs))
    batch.close()
    n = len(items)
    log(f"batch mix: plain {1000 * plain_s:.1f} ms, batched {1000 * batch_s:.1f} ms for {n} proofs "
        f"(graph replay {graph}; eager launches {1000 * eager_s:.1f} ms); bit-exact={exact} verified={verified}")
    return {"proofs": n, "shapes": shapes, "per_shape": per_shape,
            "plain_ms": round(1000 * plain_s, 3), "batched_ms": round(1000 * batch_s, 3),
            "batched_graph_replay": graph, "batched_eager_ms": round(1000 * eager_s, 3),
            "plain_proofs_per_s": round(n / plain_s, 2), "batched_proofs_per_s": round(n / batch_s, 2),
            "speedup": round(plain_s / batch_s, 3), "bit_exact_vs_plain": exact, "all_verified": verified,
            "reps": reps}


def qualification_leg(prover, wit, rs, plain_proof, window_s: float, max_p: int) -> dict:
    """The reference harness's utilisation gate (scripts/bench/qualify.py:
    11-64): for concurrency p = 1, 2, 4, ... max_p (p proofs of the 2^20
    context together -- p = 1 the plain call, p > 1 one batch call of p
    proofs), run back to back for about window_s while a sampler reads the
    device's utilisation every 100 ms; each attempt reports m / T_max (T_max
    the slowest call), the trimmed utilisation summary and its status
    (util_unqualified below 90 %).  The best attempt is the one with the most
    proofs/s among those that reach the target."""
    pci = N.device_pci_bus_id(0)
    src_name, src = monitor.source_for(pci)
    r, s_ = rs[0], rs[1]
    attempts = []
    p = 1
    batch = N.Batch()
    while True:
        if p == 1:
            def call():
                return [prover.prove_blind(wit, r, s_)]
        else:
            args = [(prover, wit, r, s_)] * p

            def call(args=args):
                return batch.prove(args)
        outs = call()  # warm-up: work buffers grow here
        ok = all(np.array_equal(o[1], plain_proof) for o in outs)
        sampler = monitor.Sampler(src) if src else None
        if sampler:
            sampler.start()
        times = []
        t_start = time.time()
        while len(times) < 3 or time.time() - t_start < window_s:
            t0 = time.perf_counter()
            outs = call()
            times.append(1000.0 * (time.perf_counter() - t0))
        t_end = time.time()
        samples = sampler.stop() if sampler else []
        ok = ok and all(np.array_equal(o[1], plain_proof) for o in outs)
        rec = monitor.qualify(times, samples, (t_start, t_end), m=p, ok=ok)
        rec["concurrency"] = p
        rec["calls"] = len(times)
        if samples:
            pw = [x.power_w for x in samples if x.power_w is not None]
            ck = [x.clock_mhz for x in samples if x.clock_mhz is not None]
            rec["power_w_mean"] = round(statistics.fmean(pw), 1) if pw else None
            rec["gfx_clock_mhz_mean"] = round(statistics.fmean(ck), 1) if ck else None
        attempts.append(rec)
        log(f"qualification p={p}: {rec['throughput_proofs_s']} proofs/s (m/T_max), util mean "
            f"{rec['util']['util_mean']} over {rec['util']['sample_count']} samples, {rec['status']}")
        if p >= max_p:
            break
        p = min(2 * p, max_p)
    batch.close()
    qualified = [a for a in attempts if a["status"] == "ok"]
    if qualified:
        best = dict(max(qualified, key=lambda a: a["throughput_proofs_s"] or 0.0), qualified=True)
    else:
        best = dict(max(attempts, key=lambda a: (a["util"]["util_mean"] or -1.0, a["concurrency"])),
                    qualified=False)
    return {"source": src_name, "pci_bus_id": pci, "sample_ms": int(1000 * monitor.SAMPLE_S),
            "target": monitor.TARGET, "trim_ratio": monitor.TRIM_RATIO,
            "best": {k: best[k] for k in ("concurrency", "throughput_proofs_s", "t_max_ms", "status", "qualified")}
            | {"util_mean": best["util"]["util_mean"]},
            "attempts": attempts}


def serving_leg(items, big, qps: float, n_req: int, verify_all: bool) -> dict:
    """A Poisson request stream at `qps` drawn from the mix (90%) and the 2^20
    matmul context (10%), served twice with the same requests: by the plain
    one-at-a-time path (1 worker, batches of 1) and by continuous batching
    (3 workers, batches up to 32, 64 GiB of admitted work memory), each
    server warmed first (every worker proves the stream's largest request
    once).  Reports p50/p99 submit-to-completion latency and proofs/s; the
    batched results must equal the plain ones bit for bit (and verify)."""
    rng = np.random.default_rng(SEED ^ 0x5E5)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n_req))
    picks = []
    for i in range(n_req):
        if rng.random() < 0.1:
            pv, keys, circ, w = big
            name = "matmul100_sparse"
        else:
            name, circ, pv, keys, w, _, _ = items[int(rng.integers(len(items)))]
        rs = circuits.random_scalars(2, (SEED << 4) ^ i)
        picks.append((name, circ, pv, keys, w, rs[0].copy(), rs[1].copy()))

    largest = max(picks, key=lambda p: N.Server.request_bytes(p[2]))

    def run(workers: int, max_batch: int):
        srv = N.Server(workers=workers, max_batch=max_batch, budget_bytes=64 << 30)
        # warm-up before traffic: every worker proves the stream's largest
        # request once (nkr_g16_server_warm), then the stream's first requests
        # go through the queue
        _, _, pv, _, w, r, s = largest
        srv.warm([(pv, w, r, s)])
        warm = [srv.submit(pv, w, r, s) for (_, _, pv, _, w, r, s) in picks[:max(1, min(8, n_req))]]
        for t in warm:
            srv.wait(t)
        tickets = []
        t0 = time.perf_counter()
        for (name, circ, pv, keys, w, r, s), at in zip(picks, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(pv, w, r, s))
        outs = [srv.wait(t) for t in tickets]
        elapsed = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        lat = [o[2] for o in outs]
        return outs, {"proofs_per_s": round(n_req / elapsed, 2), "p50_ms": round(percentile(lat, 50), 3),
                      "p99_ms": round(percentile(lat, 99), 3), "elapsed_s": round(elapsed, 3),
                      "batches_incl_warmup": int(st["batches"]), "max_batch": int(st["max_batch"])}

    log(f"serving: {n_req} requests at {qps:g}/s (Poisson), plain one-at-a-time")
    plain_outs, plain = run(1, 1)
    log(f"serving: plain p50 {plain['p50_ms']} ms p99 {plain['p99_ms']} ms, {plain['proofs_per_s']}/s; "
        "continuous batching")
    cb_outs, cb = run(3, 32)
    log(f"serving: batched p50 {cb['p50_ms']} ms p99 {cb['p99_ms']} ms, {cb['proofs_per_s']}/s")
    exact = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(plain_outs, cb_outs))
    todo = [(p, o) for p, o in zip(picks, cb_outs) if verify_all or p[0] != "matmul100_sparse"]
    verified = all(N.verify(keys, w[1:1 + circ.public_inputs], o[1]) for (_, circ, _, keys, w, _, _), o in todo)
    big_share = sum(1 for p in picks if p[0] == "matmul100_sparse")
    return {"qps": qps, "requests": n_req, "requests_2p20": big_share,
            "plain": plain, "continuous_batching": cb,
            "throughput_ratio": round(cb["proofs_per_s"] / plain["proofs_per_s"], 3),
            "bit_exact_vs_plain": exact, "verified": verified, "verified_count": len(todo)}


if __name__ == "__main__":
    sys.exit(main())

"""Parity of the HIP operators against the CPU oracle and the golden vectors,
called through the C ABI.  Integer field work: the bar is bit-exact equality."""
import ctypes

import numpy as np
import pytest

import goldens
import nkr_g16 as N
import pyoracle as O

pytestmark = pytest.mark.gpu
G = goldens.load("bn254_golden.json")


def rand_fr(seed, n):
    return O.fr_random(seed, n)


def rand_fq(seed, n):
    # uniform < q: libff rule against q, via the Fr sampler's stream then reduce
    x = O.fr_random(seed, n)
    return x  # every value < r < q is a valid Fq residue too


# ---------------------------------------------------------------- field
@pytest.mark.parametrize("field", ["fr", "fq"])
def test_field_golden(field):
    d = G[field]
    a, b = goldens.fe(d["a"]), goldens.fe(d["b"])
    assert np.array_equal(N.field_op(f"{field}_mul", a, b), goldens.fe(d["mul"]))
    assert np.array_equal(N.field_op(f"{field}_add", a, b), goldens.fe(d["add"]))
    assert np.array_equal(N.field_op(f"{field}_sub", a, b), goldens.fe(d["sub"]))
    if field == "fr":
        assert np.array_equal(N.field_op("fr_inv", a), goldens.fe(d["inv"]))


def test_fq2_golden():
    d = G["fq2"]
    a = np.array([goldens.fe(x).reshape(-1) for x in d["a"]])
    b = np.array([goldens.fe(x).reshape(-1) for x in d["b"]])
    assert np.array_equal(N.field_op("fq2_mul", a, b), np.array([goldens.fe(x).reshape(-1) for x in d["mul"]]))
    assert np.array_equal(N.field_op("fq2_sqr", a), np.array([goldens.fe(x).reshape(-1) for x in d["sqr"]]))


def test_field_random_large():
    n = 1 << 18
    a, b = rand_fr(101, n), rand_fr(102, n)
    assert np.array_equal(N.field_op("fr_mul", a, b), O.fr_mul(a, b))
    assert np.array_equal(N.field_op("fq_mul", a, b), O.fq_mul(a, b))
    assert np.array_equal(N.field_op("fr_sub", a, b), O.fr_sub(a, b))
    a2, b2 = a[: n // 2].reshape(-1, 8), b[: n // 2].reshape(-1, 8)
    assert np.array_equal(N.field_op("fq2_mul", a2, b2), O.fq2_mul(a2, b2))


# ------------------------------------------------------------------ NTT
@pytest.mark.parametrize("kind,key", [(N.FFT, "fft"), (N.IFFT, "ifft"), (N.COSET_FFT, "coset_fft"),
                                      (N.ICOSET_FFT, "icoset_fft")])
def test_ntt_golden(kind, key):
    for case in G["fft"]:
        got = N.ntt(goldens.fe(case["a"]), kind)
        assert np.array_equal(got, goldens.fe(case[key])), (key, case["n"])


@pytest.mark.parametrize("log_n", [11, 12, 13, 16, 20])
def test_ntt_vs_oracle(log_n):
    a = rand_fr(200 + log_n, 1 << log_n)
    for kind in (N.FFT, N.IFFT, N.COSET_FFT, N.ICOSET_FFT):
        assert np.array_equal(N.ntt(a, kind), O.fft(a, kind)), (log_n, kind)


def test_ntt_round_trip_2_24():
    n = 1 << 24
    a = rand_fr(224, n)
    f = N.ntt(a, N.FFT)
    assert np.array_equal(N.ntt(f, N.IFFT), a)
    c = N.ntt(a, N.COSET_FFT)
    assert np.array_equal(N.ntt(c, N.ICOSET_FFT), a)
    # FFT(delta_0) = all ones; FFT is linear
    one = O.int_to_words(O.to_mont_int(1, O.R))
    d = np.zeros((n, 4), dtype=np.uint64)
    d[0] = one
    ones = N.ntt(d, N.FFT)
    assert (ones == np.array(one, dtype=np.uint64)).all()


def test_ntt_2_24_monomial_pins_omega():
    """The transforms of x at 2^24 pin the root libfqfft uses: FFT(x)[i] =
    omega^i with omega = root_of_unity^(2^(28-24)) (field_utils.tcc:38-51; the
    2^28-th root from alt_bn128_init.cpp:60), cosetFFT(x)[i] = 5 omega^i, and
    the inverse transforms of those vectors return x.  A transform on another
    primitive root, or with permuted twiddles, fails the ratio check."""
    n, log_n = 1 << 24, 24
    root28 = 19103219067921713944291392827692070036145651957329286315305642004821462161904
    assert pow(root28, 1 << 27, O.R) == O.R - 1  # order exactly 2^28
    omega = pow(root28, 1 << (28 - log_n), O.R)
    mont = lambda v: np.array(O.int_to_words(O.to_mont_int(v, O.R)), dtype=np.uint64)
    x = np.zeros((n, 4), dtype=np.uint64)
    x[1] = mont(1)
    f = N.ntt(x, N.FFT)
    assert np.array_equal(f[0], mont(1)) and np.array_equal(f[1], mont(omega))
    # f[i + 1] = f[i] * omega for every i (and f wraps: f[n-1] * omega = 1)
    assert np.array_equal(O.fr_mul(f, np.broadcast_to(mont(omega), f.shape).copy()), np.roll(f, -1, axis=0))
    rng = np.random.default_rng(24)
    for i in rng.integers(0, n, 16).tolist() + [n // 2, n - 1]:
        assert np.array_equal(f[i], mont(pow(omega, i, O.R))), i
    c = N.ntt(x, N.COSET_FFT)
    assert np.array_equal(c, O.fr_mul(f, np.broadcast_to(mont(5), f.shape).copy()))
    assert np.array_equal(c[3], mont(5 * pow(omega, 3, O.R) % O.R))
    assert np.array_equal(N.ntt(f, N.IFFT), x)
    assert np.array_equal(N.ntt(c, N.ICOSET_FFT), x)


def test_quotient_golden_and_oracle():
    for case in G["quotient"]:
        h = N.quotient(goldens.fe(case["az"]), goldens.fe(case["bz"]), goldens.fe(case["cz"]))
        assert np.array_equal(h, goldens.fe(case["h"])), case["n"]
    n = 1 << 14
    az, bz, cz = rand_fr(301, n), rand_fr(302, n), rand_fr(303, n)
    assert np.array_equal(N.quotient(az, bz, cz), O.quotient(az, bz, cz))


# ----------------------------------------------------------------- SpMV
def test_spmv_golden():
    d = G["spmv"]
    y = N.spmv(d["rows"], d["cols"], np.array(d["row_ptr"]), np.array(d["col_idx"]),
               goldens.fe(d["values"]), goldens.fe(d["z"]))
    assert np.array_equal(y, goldens.fe(d["y"]))


def test_spmv_random_large():
    rows, cols = 1 << 16, 1 << 16
    rng = np.random.default_rng(7)
    per_row = rng.integers(0, 6, rows)
    row_ptr = np.concatenate([[0], np.cumsum(per_row)]).astype(np.uint32)
    nnz = int(row_ptr[-1])
    col_idx = rng.integers(0, cols, nnz).astype(np.uint32)
    vals = rand_fr(401, nnz)
    z = rand_fr(402, cols)
    got = N.spmv(rows, cols, row_ptr, col_idx, vals, z)
    assert np.array_equal(got, O.spmv(rows, cols, row_ptr, col_idx, vals, z))


def test_spmv_long_rows():
    # rows above 16 non-zeros go to one wave each: mix them with short ones,
    # the 16/17 boundary, empty rows and rows longer than a wave
    rows, cols = 4096, 3000
    rng = np.random.default_rng(11)
    per_row = rng.integers(0, 5, rows)
    per_row[::97] = rng.integers(17, 300, per_row[::97].size)
    per_row[5], per_row[6], per_row[7], per_row[8] = 16, 17, 0, 64
    per_row[9], per_row[10] = 65, 1000
    row_ptr = np.concatenate([[0], np.cumsum(per_row)]).astype(np.uint32)
    nnz = int(row_ptr[-1])
    col_idx = rng.integers(0, cols, nnz).astype(np.uint32)
    vals = rand_fr(403, nnz)
    z = rand_fr(404, cols)
    got = N.spmv(rows, cols, row_ptr, col_idx, vals, z)
    assert np.array_equal(got, O.spmv(rows, cols, row_ptr, col_idx, vals, z))


# ------------------------------------------------------------------ MSM
def test_fixed_base_mul_matches_oracle():
    k = rand_fr(501, 300)
    k[0] = 0
    k[1] = O.int_to_words(O.to_mont_int(1, O.R))
    g1, g2 = O.g1_generator(), O.g2_generator()
    assert np.array_equal(N.fixed_base_mul(g1, k), O.g1_to_affine(O.g1_mul(g1, k)))
    assert np.array_equal(N.fixed_base_mul(g2, k[:40], g2=True), O.g2_to_affine(O.g2_mul(g2, k[:40])))


def test_msm_golden_fixtures():
    for case in G["msm"]:
        bases = goldens.points(case["bases"])
        sc = goldens.fe(case["scalars"])
        g2 = case["group"] == "g2"
        got = N.msm(bases, sc, g2=g2)
        assert np.array_equal(got, goldens.point(case["result"])), (case["group"], case["n"], case["flavour"])


@pytest.mark.parametrize("n,flavour", [(1, "uniform"), (3, "uniform"), (4096, "uniform"), (4096, "witness"),
                                       (5000, "zeros")])
def test_msm_g1_vs_oracle(n, flavour):
    g1 = O.g1_generator()
    bases = N.fixed_base_mul(g1, rand_fr(600 + n, n))
    s = rand_fr(700 + n, n)
    if flavour == "witness":
        rng = np.random.default_rng(n)
        pick = rng.random(n) < 0.5
        bit = rng.random(n) < 0.5
        one = np.array(O.int_to_words(O.to_mont_int(1, O.R)), dtype=np.uint64)
        s[pick & bit] = one
        s[pick & ~bit] = 0
    if flavour == "zeros":
        s[:] = 0
        s[17] = O.int_to_words(O.to_mont_int(5, O.R))
    got = N.msm(bases, s)
    want = O.g1_to_affine(O.msm_g1(bases, s))[0]
    assert np.array_equal(got, want)


def test_msm_duplicate_and_opposite_bases():
    """P + P and P + (-P) inside one bucket (the reference HIP kernel's
    incomplete addition returns infinity for the former)."""
    g1 = O.g1_generator()
    base = N.fixed_base_mul(g1, rand_fr(801, 1))[0]
    neg = O.g1_to_affine(O.g1_neg(base))[0]
    bases = np.stack([base] * 40 + [neg] * 7)
    s = np.repeat(rand_fr(802, 1), 47, axis=0)
    got = N.msm(bases, s)
    assert np.array_equal(got, O.g1_to_affine(O.msm_g1(bases, s))[0])


def test_msm_g2_duplicate_and_opposite_bases():
    """The same over G2: P + P and P + (-P) inside the bucket pieces and
    between their partial sums (the reduced-radix full addition's doubling
    and cancellation)."""
    g2 = O.g2_generator()
    base = N.fixed_base_mul(g2, rand_fr(811, 1), g2=True)[0]
    minus_one = np.array(O.int_to_words(O.to_mont_int(O.R - 1, O.R)), dtype=np.uint64).reshape(1, 4)
    neg = O.g2_to_affine(O.g2_mul(base, minus_one))[0]
    bases = np.stack([base] * 40 + [neg] * 7)
    s = np.repeat(rand_fr(812, 1), 47, axis=0)
    got = N.msm(bases, s, g2=True)
    assert np.array_equal(got, O.g2_to_affine(O.msm_g2(bases, s))[0])


def test_msm_g2_vs_oracle():
    n = 2048
    g2 = O.g2_generator()
    bases = N.fixed_base_mul(g2, rand_fr(901, n), g2=True)
    s = rand_fr(902, n)
    got = N.msm(bases, s, g2=True)
    assert np.array_equal(got, O.g2_to_affine(O.msm_g2(bases, s))[0])


@pytest.mark.parametrize("log_n", [16, 18])
def test_msm_closed_form(log_n):
    """Bases k_i G: sum s_i (k_i G) = (sum k_i s_i) G, both scalar kinds."""
    n = 1 << log_n
    k = rand_fr(1000 + log_n, n)
    s = rand_fr(1100 + log_n, n)
    g1 = O.g1_generator()
    bases = N.fixed_base_mul(g1, k)
    m = N.Msm(bases)
    kk = O.fr_ints(k)
    for variant in ("uniform", "witness"):
        if variant == "witness":
            rng = np.random.default_rng(log_n)
            pick = rng.random(n) < 0.5
            bit = rng.random(n) < 0.5
            s[pick & bit] = O.int_to_words(O.to_mont_int(1, O.R))
            s[pick & ~bit] = 0
        tot = sum(a * b for a, b in zip(kk, O.fr_ints(s))) % O.R
        want = O.g1_to_affine(O.g1_mul(g1, O.fr_array([tot])))[0]
        assert np.array_equal(m.run(s), want), variant
    m.close()


# ------------------------------------------- operator sizes: closed forms
_RINV = pow(O.MONT, -1, O.R)


def block_sums(s, block):
    """Canonical sum of each run of `block` consecutive Montgomery-form
    scalars.  Montgomery form is linear (s_i R mod r), so a run's sum is R^-1
    times the sum of its words, mod r; the words are summed exactly in numpy
    as 32-bit halves (under 2^32 terms per run, every column stays below 2^64)."""
    a = np.ascontiguousarray(s, dtype=np.uint64).reshape(-1, block, 4)
    lo = (a & np.uint64(0xFFFFFFFF)).sum(axis=1, dtype=np.uint64).tolist()
    hi = (a >> np.uint64(32)).sum(axis=1, dtype=np.uint64).tolist()
    out = []
    for lr, hr in zip(lo, hi):
        t = sum((lv << (64 * w)) + (hv << (64 * w + 32)) for w, (lv, hv) in enumerate(zip(lr, hr)))
        out.append(t * _RINV % O.R)
    return out


def check_block_sums(s, t, block, runs=4):
    f = O.fr_ints(s[: runs * block])
    assert t[:runs] == [sum(f[b * block:(b + 1) * block]) % O.R for b in range(runs)]


def witness_like(s, seed):
    """Half of the scalars set to 0 or 1 (witness-like mix), the rest kept."""
    rng = np.random.default_rng(seed)
    pick = rng.random(s.shape[0]) < 0.5
    bit = rng.random(s.shape[0]) < 0.5
    s[pick & bit] = O.int_to_words(O.to_mont_int(1, O.R))
    s[pick & ~bit] = 0
    return s


def dot_closed_form(k, s):
    """sum_i k_i s_i mod r for Montgomery-form vectors: the products by the
    oracle's Montgomery multiplication (Mont(k_i s_i)), summed exactly by
    block_sums; the first 256 terms are re-derived with Python integers."""
    prod = O.fr_mul(k, s)
    head = sum(a * b for a, b in zip(O.fr_ints(k[:256]), O.fr_ints(s[:256]))) % O.R
    assert block_sums(prod[:256], 256)[0] == head
    return block_sums(prod, prod.shape[0])[0]


@pytest.mark.parametrize("log_n", [20, 24])
def test_msm_g1_operator_sizes_closed_form(log_n):
    """G1 MSM at 2^20 and 2^24 distinct bases k_i G (c = 16, 16 windows: 2^28
    bucket entries and 16 GiB of window-shifted bases at 2^24), uniform and
    witness-like scalars: sum_i s_i (k_i G) = (sum_i k_i s_i) G."""
    n = 1 << log_n
    g1 = O.g1_generator()
    k = rand_fr(1200 + log_n, n)
    m = N.Msm(N.fixed_base_mul(g1, k))
    s = rand_fr(1300 + log_n, n)
    for variant in ("uniform", "witness"):
        if variant == "witness":
            s = witness_like(s, 1400 + log_n)
        want = O.g1_to_affine(O.g1_mul(g1, O.fr_array([dot_closed_form(k, s)])))[0]
        assert np.array_equal(m.run(s), want), (log_n, variant)
    m.close()


def test_msm_g1_repeated_bases_closed_form():
    """2^20 terms whose bases repeat over runs of 64 (k_j G for run j): the
    repeated base drives the doubling and cancelling cases of the bucket
    additions; sum_i s_i B_i = (sum_j k_j T_j) G with T_j the run sums."""
    n, block = 1 << 20, 64
    g1 = O.g1_generator()
    k = rand_fr(1220, n // block)
    bases = np.repeat(N.fixed_base_mul(g1, k), block, axis=0)
    m = N.Msm(bases)
    del bases
    kk = O.fr_ints(k)
    s = witness_like(rand_fr(1320, n), 1420)
    t = block_sums(s, block)
    check_block_sums(s, t, block)
    tot = sum(a * b for a, b in zip(kk, t)) % O.R
    want = O.g1_to_affine(O.g1_mul(g1, O.fr_array([tot])))[0]
    assert np.array_equal(m.run(s), want)
    m.close()


def test_msm_g2_2_18_closed_form():
    """G2 MSM at 2^18 distinct bases k_i G2 with uniform scalars."""
    n = 1 << 18
    g2 = O.g2_generator()
    k = rand_fr(1500, n)
    bases = N.fixed_base_mul(g2, k, g2=True)
    s = rand_fr(1501, n)
    want = O.g2_to_affine(O.g2_mul(g2, O.fr_array([dot_closed_form(k, s)])))[0]
    assert np.array_equal(N.msm(bases, s, g2=True), want)


def test_spmv_2_20_rows():
    """CSR SpMV at 2^20 rows, 3 non-zeros per row on average (Poisson), uniform
    Fr coefficients, against the oracle."""
    rows = cols = 1 << 20
    rng = np.random.default_rng(13)
    per_row = rng.poisson(3, rows)
    row_ptr = np.concatenate([[0], np.cumsum(per_row)]).astype(np.uint32)
    nnz = int(row_ptr[-1])
    col_idx = rng.integers(0, cols, nnz).astype(np.uint32)
    vals = rand_fr(1601, nnz)
    z = rand_fr(1602, cols)
    got = N.spmv(rows, cols, row_ptr, col_idx, vals, z)
    assert np.array_equal(got, O.spmv(rows, cols, row_ptr, col_idx, vals, z))


def test_spmv_context_device_vectors():
    """nkr_spmv_create / nkr_spmv_run_dev (spMVMContext): the matrix uploaded
    once, two different device vectors, each equal to the oracle; bad CSR
    input is rejected at create."""
    rows, cols = 5000, 3000
    rng = np.random.default_rng(17)
    per_row = rng.integers(0, 6, rows)
    per_row[::101] = 40  # wave-per-row path
    row_ptr = np.concatenate([[0], np.cumsum(per_row)]).astype(np.uint32)
    nnz = int(row_ptr[-1])
    col_idx = rng.integers(0, cols, nnz).astype(np.uint32)
    vals = rand_fr(1701, nnz)
    sp = N.Spmv(rows, cols, row_ptr, col_idx, vals)
    lib = N.lib()
    ydev = ctypes.c_void_p()
    N._check(lib.nkr_dev_alloc(rows * 32, ctypes.byref(ydev)))
    for seed in (1702, 1703):
        z = rand_fr(seed, cols)
        zdev = ctypes.c_void_p()
        N._check(lib.nkr_dev_alloc(z.nbytes, ctypes.byref(zdev)))
        N._check(lib.nkr_memcpy_htod(zdev, N._p(z), z.nbytes))
        sp.run_dev(zdev.value, ydev.value)
        N._check(lib.nkr_device_synchronize())
        y = np.zeros((rows, 4), dtype=np.uint64)
        N._check(lib.nkr_memcpy_dtoh(N._p(y), ydev, y.nbytes))
        N._check(lib.nkr_dev_free(zdev))
        assert np.array_equal(y, O.spmv(rows, cols, row_ptr, col_idx, vals, z)), seed
    N._check(lib.nkr_dev_free(ydev))
    sp.close()
    bad = col_idx.copy()
    bad[7] = cols
    with pytest.raises(N.NkrError, match="out-of-range column"):
        N.Spmv(rows, cols, row_ptr, bad, vals)

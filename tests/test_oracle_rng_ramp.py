"""Pins of the oracle's xorshift64* generator and of its viscosity /
diffusivity spin-up ramp, independent of the oracle's own formulas.

xorshift64* (src/uniform/rng.rs:11-30; Vigna, "An experimental exploration
of Marsaglia's xorshift generators, scrambled", ACM TOMS 42(4), 2016: the
state update of Marsaglia's shift triple (12, 25, 27) followed by the
multiplication by M32 = 2685821657736338717 modulo 2^64):
  - the state update is probed column by column as a 64x64 matrix over GF(2);
    it must be linear and have order exactly 2^64 - 1 (a full period, which
    only a few thousand of the possible shift triples give), and equal the
    published triple's matrix built from shift matrices;
  - the scrambler is recovered from (state, output) pairs with the modular
    inverse of the state and must be the published constant;
  - next_f64 * 2^53 must be the integer next_u64 >> 11;
  - the initial temperature noise is rebuilt bitwise from a generator written
    here in matrix form (seed | 1, (iy, ix) draw order).

Ramp (src/uniform/solver.rs:390-410, :433-442), through orc_step_scalars:
closed forms nu_eff(0) = m nu, nu_eff(S/2) = sqrt(m) nu, nu_eff(>= S) = nu,
log-linearity in the step, the same for alpha, beta(S/2) = beta_eff / 2, and
the 0.002/3 floor of the molecular tau.

Each group of checks is also run on implementations with a plausible mistake
planted (a wrong or mis-directed shift, a dropped or wrong multiplier, a
halved float conversion, a missing seed bit; an inverted or linear ramp, an
unramped alpha, an off-by-one beta, a missing tau floor): every one of them
must fail at least one check while the oracle passes them all.
"""
import ctypes as C
import math

import numpy as np
import pytest

from tests.helpers import SMALL, args_of, bits_equal

MASK = (1 << 64) - 1
M32 = 2685821657736338717          # Vigna 2016, xorshift64* multiplier
TRIPLE = (12, 25, 27)              # Marsaglia's shift triple used with it
PRIMES_2_64_M1 = (3, 5, 17, 257, 641, 65537, 6700417)


# ---------------------------------------------------------------------------
# GF(2) linear algebra on 64-bit words

def bits_of(x):
    return np.array([(x >> i) & 1 for i in range(64)], dtype=np.int64)


def word_of(b):
    return sum(int(v) << i for i, v in enumerate(b) if v)


def matrix_of(update):
    """Column j is the image of the basis vector e_j."""
    m = np.zeros((64, 64), dtype=np.int64)
    for j in range(64):
        m[:, j] = bits_of(update(1 << j))
    return m


def shift_matrix(k):
    """x -> x >> k (k > 0) or x << -k (k < 0), as a GF(2) matrix."""
    m = np.zeros((64, 64), dtype=np.int64)
    for j in range(64):
        i = j - k
        if 0 <= i < 64:
            m[i, j] = 1
    return m


def mat_mul(a, b):
    return (a @ b) & 1


def mat_pow(m, e):
    r = np.eye(64, dtype=np.int64)
    while e:
        if e & 1:
            r = mat_mul(r, m)
        m = mat_mul(m, m)
        e >>= 1
    return r


def published_matrix():
    """x ^= x >> a; x ^= x << b; x ^= x >> c  =  (I + R_c)(I + L_b)(I + R_a)."""
    a, b, c = TRIPLE
    eye = np.eye(64, dtype=np.int64)
    ta = (eye + shift_matrix(a)) & 1
    tb = (eye + shift_matrix(-b)) & 1
    tc = (eye + shift_matrix(c)) & 1
    return mat_mul(tc, mat_mul(tb, ta))


# ---------------------------------------------------------------------------
# implementations under test: the oracle and planted mistakes

class OracleRng:
    """orc_rng_next_u64 / orc_rng_next_f64 on a caller-held state."""

    def __init__(self, oracle):
        self.L = oracle.lib()

    def next_u64(self, state):
        st = C.c_uint64(state)
        out = self.L.orc_rng_next_u64(C.byref(st))
        return st.value, out

    def next_f64(self, state):
        st = C.c_uint64(state)
        out = self.L.orc_rng_next_f64(C.byref(st))
        return st.value, out

    def seed_state(self, seed):
        return seed | 1


class PlantedRng:
    def __init__(self, shifts=((12, ">"), (25, "<"), (27, ">")), mult=M32, f64_shift=11,
                 seed_or=1):
        self.shifts, self.mult, self.f64_shift, self.seed_or = shifts, mult, f64_shift, seed_or

    def next_u64(self, state):
        x = state
        for k, d in self.shifts:
            x ^= (x >> k) if d == ">" else ((x << k) & MASK)
        return x, (x * self.mult) & MASK

    def next_f64(self, state):
        st, out = self.next_u64(state)
        return st, (out >> self.f64_shift) / float(1 << (64 - self.f64_shift))

    def seed_state(self, seed):
        return seed | self.seed_or


PLANTED_RNG = {
    "triple_off_by_one": PlantedRng(shifts=((12, ">"), (25, "<"), (26, ">"))),
    "other_full_period_triple": PlantedRng(shifts=((13, ">"), (7, "<"), (17, ">"))),
    "second_shift_flipped": PlantedRng(shifts=((12, ">"), (25, ">"), (27, ">"))),
    "multiplier_dropped": PlantedRng(mult=1),
    "multiplier_wrong_digit": PlantedRng(mult=M32 ^ 0x2),
    "f64_from_52_bits": PlantedRng(f64_shift=12),
    "seed_not_forced_odd": PlantedRng(seed_or=0),
}


# ---------------------------------------------------------------------------
# the checks (each returns a list of failures; empty = pass)

def rng_checks(impl, rs):
    fails = []
    update = lambda s: impl.next_u64(s)[0]  # noqa: E731
    m = matrix_of(update)
    # linear over GF(2)
    for _ in range(32):
        a, b = int(rs.integers(0, 1 << 63)) * 2 + 1, int(rs.integers(0, 1 << 63))
        if update(a ^ b) != update(a) ^ update(b):
            fails.append("state update is not linear over GF(2)")
            break
    # full period: order of m is exactly 2^64 - 1
    eye = np.eye(64, dtype=np.int64)
    if not np.array_equal(mat_pow(m, MASK), eye):
        fails.append("m^(2^64-1) != I")
    else:
        for p in PRIMES_2_64_M1:
            if np.array_equal(mat_pow(m, MASK // p), eye):
                fails.append(f"period divides (2^64-1)/{p}")
                break
    if not np.array_equal(m, published_matrix()):
        fails.append("state update is not the (12,25,27) triple")
    # scrambler: out = state * M32 mod 2^64
    mults = set()
    s = 0x9E3779B97F4A7C15
    for _ in range(64):
        s, out = impl.next_u64(s)
        if s & 1:
            mults.add((out * pow(s, -1, 1 << 64)) & MASK)
    if mults != {M32}:
        fails.append(f"multiplier {sorted(mults)[:2]} != M32")
    # next_f64 = (next_u64 >> 11) / 2^53
    s = 12345
    for _ in range(64):
        s2, out = impl.next_u64(s)
        s3, v = impl.next_f64(s)
        if s3 != s2 or not (0.0 <= v < 1.0) or v * 2.0 ** 53 != float(out >> 11):
            fails.append("next_f64 != (next_u64 >> 11) / 2^53")
            break
        s = s2
    return fails


def init_noise_matches(impl, oracle_solver, nx, ny, seed):
    """Rebuild the initial temperature (solver.rs:160-176) from `impl`."""
    st = impl.seed_state(seed)
    u = np.empty(nx * ny)
    for k in range(nx * ny):
        st, u[k] = impl.next_f64(st)
    lin = 0.5 + (-1.0) * ((np.arange(nx * ny) % nx) / (nx - 1))
    return bits_equal(oracle_solver.temperature, lin + 1.0e-3 * (2.0 * u - 1.0))


class MatrixRng:
    """A generator written here in matrix form: state_{k+1} = M s_k over
    GF(2) with M the published triple's matrix, output = state * M32."""

    def __init__(self):
        self.m = published_matrix()

    def next_u64(self, state):
        s = word_of(mat_mul(self.m, bits_of(state)[:, None])[:, 0])
        return s, (s * M32) & MASK

    def next_f64(self, state):
        s, out = self.next_u64(state)
        return s, (out >> 11) / float(1 << 53)

    def seed_state(self, seed):
        return seed | 1


def test_published_matrix_has_full_period():
    m = published_matrix()
    eye = np.eye(64, dtype=np.int64)
    assert math.prod(PRIMES_2_64_M1) == MASK
    assert np.array_equal(mat_pow(m, MASK), eye)
    for p in PRIMES_2_64_M1:
        assert not np.array_equal(mat_pow(m, MASK // p), eye)


def test_oracle_rng_passes_every_check(oracle):
    assert rng_checks(OracleRng(oracle), np.random.default_rng(1)) == []


@pytest.mark.parametrize("name", sorted(PLANTED_RNG))
def test_planted_rng_mistake_fails(oracle, name):
    impl = PLANTED_RNG[name]
    fails = rng_checks(impl, np.random.default_rng(1))
    if name == "seed_not_forced_odd":
        # the generator itself is right; the init noise rebuilt from an even
        # seed must differ from the oracle's
        assert fails == []
        o = oracle.OracleSolver(*args_of(SMALL), 1)
        assert not init_noise_matches(impl, o, SMALL["nx"], SMALL["ny"], SMALL["seed"])
    else:
        assert fails, name


def test_init_noise_from_matrix_generator(oracle):
    """The oracle's initial temperature equals the linear profile plus the
    noise of the matrix-form generator seeded with seed | 1, drawn in
    (iy outer, ix inner) order; seeds 42 (even) and 7 (odd)."""
    for seed in (42, 7):
        cfg = dict(SMALL, seed=seed)
        o = oracle.OracleSolver(*args_of(cfg), 1)
        assert init_noise_matches(MatrixRng(), o, cfg["nx"], cfg["ny"], seed)


# ---------------------------------------------------------------------------
# spin-up ramp of beta, nu and alpha (solver.rs:390-410, :433-442)

def oracle_scalars(o):
    return lambda k: o.step_scalars(k)


def planted_scalars(d, mistake):
    nu, alpha, beta_eff, S, m = (d["nu"], d["alpha"], d["beta_eff"], int(d["spin_up_steps"]),
                                 d["nu_start_mult"])

    def sc(k):
        x = k / S
        ramp = k < S
        if mistake == "inverted_exponent":
            nu_e = nu * m * (m) ** x if ramp else nu
            al_e = alpha * m * (m) ** x if ramp else alpha
        elif mistake == "linear_ramp":
            nu_e = nu * m + (nu - nu * m) * x if ramp else nu
            al_e = alpha * m + (alpha - alpha * m) * x if ramp else alpha
        else:
            nu_e = nu * m * (1.0 / m) ** x if ramp else nu
            al_e = alpha * m * (1.0 / m) ** x if ramp else alpha
        if mistake == "alpha_unramped":
            al_e = alpha
        beta = beta_eff * ((k + 1) / S if mistake == "beta_off_by_one" else x) if ramp else beta_eff
        tau = 0.5 + 3.0 * (nu_e if mistake == "no_tau_floor" else max(nu_e, 0.002 / 3.0))
        return dict(beta=beta, nu_eff=nu_e, alpha_eff=al_e, tau=tau)

    return sc


def ramp_checks(sc, d):
    nu, alpha, beta_eff, S, m = (d["nu"], d["alpha"], d["beta_eff"], int(d["spin_up_steps"]),
                                 d["nu_start_mult"])
    fails = []

    def close(a, b, tol=2e-15):
        return abs(a - b) <= tol * abs(b)

    s0 = sc(0)
    if not (s0["nu_eff"] == nu * m and s0["alpha_eff"] == alpha * m and s0["beta"] == 0.0):
        fails.append("step 0: nu_eff = m nu, alpha_eff = m alpha, beta = 0")
    assert S % 2 == 0
    sh = sc(S // 2)
    if not (close(sh["nu_eff"], nu * math.sqrt(m)) and close(sh["alpha_eff"], alpha * math.sqrt(m))):
        fails.append("step S/2: sqrt(m) nu, sqrt(m) alpha")
    if sh["beta"] != beta_eff * 0.5:
        fails.append("beta(S/2) != beta_eff / 2")
    for k in (S, S + 1, 7 * S):
        sk = sc(k)
        if not (sk["nu_eff"] == nu and sk["alpha_eff"] == alpha and sk["beta"] == beta_eff):
            fails.append(f"step {k} >= S is not the target")
            break
    # log-linear in the step: log nu_eff(k) = log(m nu) - (k/S) log m
    for k in (1, S // 3, S - 1):
        sk = sc(k)
        for key, base in (("nu_eff", nu), ("alpha_eff", alpha)):
            want = math.log(m * base) - (k / S) * math.log(m)
            if abs(math.log(sk[key]) - want) > 1e-13:
                fails.append(f"{key}({k}) not geometric")
        if sk["beta"] != beta_eff * (k / S):
            fails.append(f"beta({k}) not linear")
    # molecular tau: 0.5 + 3 nu_eff above the floor, 0.502 below it
    for k in (0, S // 2, S):
        sk = sc(k)
        if sk["nu_eff"] > 0.002 / 3.0:
            if not close((sk["tau"] - 0.5) / 3.0, sk["nu_eff"], 1e-12):
                fails.append(f"tau({k}) != 0.5 + 3 nu_eff")
        elif not close(sk["tau"], 0.502, 1e-15):
            fails.append(f"tau({k}) below the floor != 0.502")
    return fails


RAMP_CASES = [
    # (nu_start_mult, nu_override, alpha_mult): a ramp above the floor, and
    # one that ends below it (nu = 4e-4 < 0.002/3; alpha_mult 5 keeps the
    # thermal tau above the constructor's 0.505 bound)
    (4.0, None, 1.0),
    (2.5, 4.0e-4, 5.0),
]


@pytest.mark.parametrize("m,nu_override,alpha_mult", RAMP_CASES)
def test_oracle_ramp_closed_forms(oracle, m, nu_override, alpha_mult):
    o = oracle.OracleSolver(*args_of(SMALL), 1, alpha_mult=alpha_mult, nu_start_mult=m,
                            nu_override=nu_override)
    d = o.derived()
    assert d["nu_start_mult"] == m
    assert ramp_checks(oracle_scalars(o), d) == []


def test_oracle_ramp_identity_without_multiplier(oracle):
    """nu_start_mult = 1: nu_eff and alpha_eff equal nu and alpha bitwise on
    every step of the spin-up (the quirk the reference keeps)."""
    o = oracle.OracleSolver(*args_of(SMALL), 1)
    d = o.derived()
    for k in range(0, int(d["spin_up_steps"]) + 2):
        s = o.step_scalars(k)
        assert s["nu_eff"] == d["nu"] and s["alpha_eff"] == d["alpha"]


@pytest.mark.parametrize("mistake", ["inverted_exponent", "linear_ramp", "alpha_unramped",
                                     "beta_off_by_one", "no_tau_floor"])
def test_planted_ramp_mistake_fails(oracle, mistake):
    failed = False
    for m, nu_override, alpha_mult in RAMP_CASES:
        o = oracle.OracleSolver(*args_of(SMALL), 1, alpha_mult=alpha_mult, nu_start_mult=m,
                                nu_override=nu_override)
        d = o.derived()
        assert ramp_checks(planted_scalars(d, None), d) == []  # the unplanted restatement passes
        failed |= bool(ramp_checks(planted_scalars(d, mistake), d))
    assert failed, mistake

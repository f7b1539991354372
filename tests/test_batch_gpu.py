"""The batched uneven chain on the MI355X (csd_ring_batch_* through the C ABI): ciphertexts at
different RNS levels go through pb_to_crt -> CRT-basis plaintext product -> crt_to_pb ->
mod_switch_down by one limb as one work list. Every ciphertext must be bit-exact with the
oracle's per-ciphertext call sequence (ring.rs:1118-1161, bfv.rs:628-662), for eager runs,
repeated hipGraph replays, explicit launch shapes and the autotuner's choice."""
import numpy as np
import pytest

from _pkg import load
from oracle import ring as R
from test_hal_gpu import from_dev, same, to_dev
from test_ring_gpu import uniform_residues

pytestmark = pytest.mark.gpu

FIELD, LOG2_Q, T_BITS = "field_36864", (960, 980), 76


@pytest.fixture(scope="module")
def setup(cpu):
    pkg = load()
    gpu = pkg.HipBackend(0)
    g = pkg.ring.RingContext(gpu, FIELD, LOG2_Q, T_BITS)
    c = R.RingContext(cpu, R.FieldDescriptor(FIELD), LOG2_Q, T_BITS)
    assert g.primes == c.primes
    yield gpu, g, c
    g.close()
    gpu.close()


def oracle_chain(cpu, c, ct, pt):
    """pb_to_crt_tensor, hal mul_mont64 by the plaintext, crt_to_pb_tensor, mod_switch_down by one
    limb: the reference call sequence for one ciphertext at level r"""
    r = ct.shape[1]
    base = list(range(r))
    x = ct.copy()
    c.pb_to_crt_tensor(x, base)
    cpu.hal_mul_mont64(x, np.ascontiguousarray(pt[:, :r]), c.mods_for(base))
    c.crt_to_pb_tensor(x, base)
    out = R.mod_switch_down(c, x.copy(), base, r - 1)
    return x, out


def make_case(setup, cpu, limbs, seed):
    gpu, g, c = setup
    base = g.initial_rns_base()
    primes = [g.primes[i] for i in base]
    rng = np.random.default_rng(seed)
    n2, n3 = g.info["deg_lprime"], g.info["deg_k"]
    cts = [uniform_residues(rng, primes[:r], (2, r, n2, n3)) for r in limbs]
    pt = uniform_residues(rng, primes, (1, len(base), n2, n3))
    expect = [oracle_chain(cpu, c, ct, pt) for ct in cts]
    return base, cts, pt, expect


def check_run(gpu, plan, dcts, douts, cts, expect, graph):
    for d, h in zip(dcts, cts):
        gpu.hal_import(d, h)
    plan.run(graph=graph)
    for d, o, (x, out) in zip(dcts, douts, expect):
        assert same(from_dev(gpu, d), x)
        assert same(from_dev(gpu, o), out)


@pytest.fixture(scope="module")
def case(setup, cpu):
    gpu, g, c = setup
    limbs = [3, 17, 5, 9, 2, 12]
    base, cts, pt, expect = make_case(setup, cpu, limbs, 31)
    dcts = [to_dev(gpu, x) for x in cts]
    douts = [gpu.hal_allocate_i64((2, r - 1) + x.shape[2:]) for r, x in zip(limbs, cts)]
    dpt = to_dev(gpu, pt)
    plan = g.batch_plan(base, dcts, douts, dpt)
    yield plan, dcts, douts, cts, expect
    plan.close()


def test_batch_eager_parity(setup, case):
    gpu = setup[0]
    plan, dcts, douts, cts, expect = case
    info = plan.info()
    assert info["slices"] == 2 * sum(x.shape[1] for x in cts)
    assert info["elements"] == 2 * len(cts)
    check_run(gpu, plan, dcts, douts, cts, expect, graph=False)


def test_batch_graph_replays(setup, case):
    """the captured chain, replayed three times with eager uploads and reads in between"""
    gpu = setup[0]
    plan, dcts, douts, cts, expect = case
    for _ in range(3):
        check_run(gpu, plan, dcts, douts, cts, expect, graph=True)


@pytest.mark.parametrize("tune", [dict(rows=1, threads=128, col_threads=64, g2=1, g3=1, persistent=0),
                                  dict(rows=2, threads=512, col_threads=128, g2=2, g3=2, persistent=0),
                                  dict(rows=4, threads=384, col_threads=192, g2=3, g3=1, persistent=0),
                                  dict(rows=0, threads=0, col_threads=256, g2=3, g3=2, persistent=0),
                                  dict(rows=1, g2=3, g3=2, persistent=1, grid=0),
                                  dict(rows=2, g2=3, g3=2, persistent=1, grid=16),
                                  dict(rows=3, g2=3, g3=2, persistent=1, grid=7)])
def test_batch_launch_shapes(setup, case, tune):
    """every launch shape and pass grouping, and the persistent one-launch chain with grids
    smaller than the slice count (each workgroup then loops over many slices), compute the same
    residues"""
    gpu = setup[0]
    plan, dcts, douts, cts, expect = case
    plan.set_tuning(**tune)
    got = plan.tuning()
    assert all(got[k] == v for k, v in tune.items())
    check_run(gpu, plan, dcts, douts, cts, expect, graph=False)
    check_run(gpu, plan, dcts, douts, cts, expect, graph=True)


def test_batch_autotune(setup, case):
    gpu = setup[0]
    plan, dcts, douts, cts, expect = case
    chosen, ms = plan.autotune(reps=1)
    assert ms > 0
    assert plan.tuning() == chosen
    check_run(gpu, plan, dcts, douts, cts, expect, graph=True)
    plan.set_tuning(rows=0, threads=0, col_threads=256, g2=3, g3=2, persistent=0, grid=0)


def test_batch_single_and_empty(setup, cpu):
    gpu, g, c = setup
    base, cts, pt, expect = make_case(setup, cpu, [17], 5)
    dcts = [to_dev(gpu, cts[0])]
    douts = [gpu.hal_allocate_i64((2, 16, 12, 3072))]
    dpt = to_dev(gpu, pt)
    plan = g.batch_plan(base, dcts, douts, dpt)
    check_run(gpu, plan, dcts, douts, cts, expect, graph=True)
    plan.close()
    empty = g.batch_plan(base, [], [], dpt)
    empty.run(graph=True)
    empty.run(graph=False)
    assert empty.info()["slices"] == 0
    empty.close()


def test_batch_errors(setup):
    gpu, g, c = setup
    pkg = load()
    base = g.initial_rns_base()
    pt = gpu.hal_allocate_i64((1, 17, 12, 3072))
    one = gpu.hal_allocate_i64((2, 1, 12, 3072))
    with pytest.raises(pkg.HalError):
        g.batch_plan(base, [one], [gpu.hal_allocate_i64((2, 0, 12, 3072))], pt)  # a level-1 ciphertext
    short = gpu.hal_allocate_i64((1, 3, 12, 3072))
    ct = gpu.hal_allocate_i64((2, 4, 12, 3072))
    with pytest.raises(pkg.HalError):
        g.batch_plan(base, [ct], [gpu.hal_allocate_i64((2, 3, 12, 3072))], short)  # plaintext shorter than the base
    plan = g.batch_plan(base, [ct], [gpu.hal_allocate_i64((2, 3, 12, 3072))], pt)
    with pytest.raises(pkg.HalError):
        plan.set_tuning(g2=4)
    with pytest.raises(pkg.HalError):
        plan.set_tuning(rows=7)  # 7 lines of 3072 do not fit the 160 KiB LDS
    with pytest.raises(pkg.HalError):
        plan.set_tuning(rows=1, persistent=1, g2=2)  # the persistent chain runs the default grouping only
    plan.close()

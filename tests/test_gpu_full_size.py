"""Full-size (logN=16, L34) parity of the batched shapes the benchmark times.

* Batched rotations by per-entry steps (``rotate_by_steps``: one fused key
  switch with per-entry keys, per-entry Galois tables and the XCD-remapped
  grids over a real batch) at levels 0, 16 and 32: sampled entries equal the
  oracle's restatement of the reference rotation (ckks_engine.py:2183-2207,
  3822-3897 with hybrid_keyswitch.py:321-371) bit for bit, every entry equals
  its own single-entry rotation, and one entry decrypts within the
  reference's tolerance.
* Batched HMult (fused tensor-product front, relinearization seeded into the
  key switch, rescale inside ModDown) at levels 0, 16 and 31: sampled entries
  equal the oracle's multiply -> relinearize -> rescale bit for bit.
* The 128-diagonal matrix-vector product at level 0 decrypts within
  128 x 2e-5 of binary64.
* The serving loop: six requests at start levels 0, 3 and 16 give the same
  bits alone and in a mix, and decrypt within window x 2e-5.

Tolerances are the reference's per-operation bounds at scale 2^40
(tests/test_ckks_operation_correctness.py:48-51).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from fhelium_amd.linear import diagonal_matvec, prepare_diagonals
from fhelium_amd.serving import PolynomialWindowSum, Request, ServingLoop, logistic_chebyshev
from oracle.engine_ref import EngineRef
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL_KEYSWITCH = 2e-5
TOL_MULTIPLY = 3e-5
STEPS = [1, 5, 127, 512, 1024, -3, 3000, 17]


@pytest.fixture(scope="module")
def full():
    eng = CkksEngine(H.config(16), rng_seed=20260807)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    for s in STEPS:
        eng.rotation_key(s)
    return eng, EngineRef(H.host_tables(16))


def _encrypt(eng, level, batch, seed, radius=1.0):
    x = np.random.default_rng(seed).uniform(-radius, radius, (batch, eng.config.slots))
    return x, eng.encrypt(x, level=level)


def _entry(ct, b):
    return ct.data[:, b: b + 1]


@pytest.mark.parametrize("level", [0, 16, 32])
def test_rotate_by_steps_full_size(full, level):
    eng, ref = full
    x, ct = _encrypt(eng, level, len(STEPS), 300 + level)
    print(f"rotate_by_steps logN=16 level={level} B={len(STEPS)} steps={STEPS}", flush=True)
    out = eng.rotate_by_steps(ct, STEPS)
    got = H.to_np(out.data)
    src = H.to_np(ct.data)
    for b in (0, 6):
        key = eng.rotation_key(STEPS[b])
        want = ref.rotate(src[:, b: b + 1], H.to_np(key.data), key.galois_element, level)
        np.testing.assert_array_equal(got[:, b: b + 1], want, err_msg=f"entry {b} step {STEPS[b]}")
    for b, s in enumerate(STEPS):
        single = ct.clone()
        single.data = _entry(ct, b).contiguous()
        assert torch.equal(eng.rotate_by_step(single, s).data, out.data[:, b: b + 1]), f"entry {b}"
    one = out.clone()
    one.data = _entry(out, 1)
    err = np.abs(eng.decrypt(one).real[0] - np.roll(x[1], STEPS[1])).max()
    assert err < TOL_KEYSWITCH, err


@pytest.mark.parametrize("level", [0, 16, 31])
def test_batched_hmult_full_size(full, level):
    eng, ref = full
    batch = 4
    x, a = _encrypt(eng, level, batch, 400 + level)
    y, b = _encrypt(eng, level, batch, 500 + level)
    print(f"batched hmult logN=16 level={level} B={batch}", flush=True)
    out = eng.hmult(a, b)
    assert out.level == level + 1
    got = H.to_np(out.data)
    an, bn = H.to_np(a.data), H.to_np(b.data)
    rlk = H.to_np(eng.relinearization_key.data)
    for e in (0, batch - 1):
        sa, sb = an[:, e: e + 1], bn[:, e: e + 1]
        prod = ref.multiply(ref.coefficient_domain_to_ntt_domain(sa, level),
                            ref.coefficient_domain_to_ntt_domain(sb, level), level)
        want = ref.rescale(ref.relinearize(prod, rlk, level), level)
        np.testing.assert_array_equal(got[:, e: e + 1], want, err_msg=f"entry {e}")
    one = out.clone()
    one.data = _entry(out, 2)
    err = np.abs(eng.decrypt(one).real[0] - x[2] * y[2]).max()
    assert err < TOL_MULTIPLY, err


def test_matvec_128_diagonals_full_size(full):
    eng, _ = full
    k = 128
    rng = np.random.default_rng(909)
    x = rng.uniform(-1, 1, eng.config.slots)
    diags = rng.uniform(-1, 1, (k, eng.config.slots))
    print(f"matvec logN=16 level=0 diagonals={k}", flush=True)
    ct = eng.encrypt(x, level=0)
    pts = prepare_diagonals(eng, diags, 0)
    out = diagonal_matvec(eng, ct, pts)
    want = sum(diags[i] * np.roll(x, i) for i in range(k))
    assert out.level == 1
    err = np.abs(eng.decrypt(out).real - want).max()
    print(f"matvec max-abs error {err:.3e}")
    assert err < k * TOL_KEYSWITCH


def _unit_clock():
    t = [0.0]

    def clock():
        now = t[0]
        t[0] += 1.0
        return now

    def sleep(dt):
        t[0] += dt

    return clock, sleep


def test_serving_full_size(full):
    eng, _ = full
    window = 64
    prog = PolynomialWindowSum(eng, logistic_chebyshev(7), radius=4.0, window=window)
    rng = np.random.default_rng(77)
    levels = [0, 0, 16, 0, 16, 3]
    xs = [rng.uniform(-4, 4, eng.config.slots) for _ in levels]
    cts = [eng.encrypt(v, level=lvl) for v, lvl in zip(xs, levels)]

    def serve(arrivals, **kw):
        reqs = [Request(i, c.clone(), a) for i, (c, a) in enumerate(zip(cts, arrivals))]
        loop = ServingLoop(prog, **kw)
        clock, sleep = _unit_clock()
        stats = loop.run(reqs, clock=clock, sleep=sleep)
        torch.cuda.synchronize()
        return reqs, stats, loop

    print("serving logN=16: 6 requests alone", flush=True)
    alone = []
    for i in range(len(cts)):
        reqs = [Request(0, cts[i].clone(), 0.0)]
        clock, sleep = _unit_clock()
        ServingLoop(prog).run(reqs, clock=clock, sleep=sleep)
        alone.append(reqs[0].output.data.clone())
        err = np.abs(eng.decrypt(reqs[0].output).real - prog.expected(xs[i])).max()
        assert err < window * TOL_KEYSWITCH, (i, err)
    print("serving logN=16: 6 requests in a mix", flush=True)
    reqs, stats, loop = serve([0.0, 0.0, 1.5, 3.0, 3.2, 6.5], max_batch=4)
    assert stats.completed == len(cts)
    assert any(len(ids) > 1 for _, _, _, ids in loop.trace)
    for r in reqs:
        assert torch.equal(r.output.data, alone[r.rid]), f"request {r.rid} differs in the mix"
        assert r.output.level == levels[r.rid] + prog.depth


@pytest.mark.parametrize("level", [0, 16, 32])
def test_hoisted_rotations_full_size(full, level):
    """The fused hoisted key switch (one decomposition of c1 shared by every
    key, each key's MAC reading it through its NTT-domain Galois map) at
    logN=16: sampled entries equal the oracle's hoisted path
    (hybrid_keyswitch.py:373-504) bit for bit, every entry equals the
    per-op hoisted sequence, and one entry decrypts within tolerance."""
    eng, ref = full
    x, ct = _encrypt(eng, level, 1, 700 + level)
    ct.data = ct.data[:, 0]
    print(f"hoisted rotations logN=16 level={level} keys={len(STEPS)}", flush=True)
    out = eng.rotate_hoisted(ct, STEPS)
    got = H.to_np(out.data)
    src = H.to_np(ct.data)[:, None]
    prepared = ref.prepare_rotation_digits(src[1], level)
    for e in (0, 5):
        key = eng.rotation_key(STEPS[e])
        want = ref.rotate_hoisted(src, prepared, H.to_np(key.data), key.galois_element, level)
        np.testing.assert_array_equal(got[:, e: e + 1], want, err_msg=f"entry {e} step {STEPS[e]}")
    eng.fused_keyswitch = False
    try:
        unfused = eng.rotate_hoisted(ct, STEPS)
    finally:
        eng.fused_keyswitch = True
    assert torch.equal(unfused.data, out.data)
    one = out.clone()
    one.data = out.data[:, 2]
    err = np.abs(eng.decrypt(one).real - np.roll(x[0], STEPS[2])).max()
    assert err < TOL_KEYSWITCH, err

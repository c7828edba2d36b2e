"""CkksEngine on gfx950: bit-exact parity with the oracle's restatement of
the reference engine methods, and decrypted error against an FP64 cleartext.

Parity (logN=12 over the full L34 chain, 39 QP rows, identical keys and
inputs; and at logN=16 for the fused HMult and rotation at levels 0 and
30): multiply, relinearize, rescale_to_next_level, rotate_by_step and
hoisted rotations reproduce the reference CPU residues exactly; the NTT
conversion is exact in exact mode and equal modulo q (canonical) in fast
mode.  Error (logN=16, the L34 preset): decrypted max-abs error of
HMult / rotation vs binary64 stays within the reference's tolerance.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from fhelium_amd.core import MaximumLevelError, CkksStateError
from oracle.engine_ref import EngineRef
from tests import helpers as H

pytestmark = pytest.mark.gpu

# reference tolerances at scale 2^40 (tests/test_ckks_operation_correctness.py:48-51)
TOL_KEYSWITCH = 2e-5
TOL_MULTIPLY = 3e-5


@pytest.fixture(scope="module", params=["fast", "exact"])
def small(request):
    cfg = H.config(12)
    eng = CkksEngine(cfg, ntt_mode=request.param, rng_seed=7)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    return eng, EngineRef(H.host_tables(12)), request.param


def _np(t):
    return H.to_np(t)


def _fresh(eng, level, batch=(2,), seed=0):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, size=batch + (eng.config.slots,))
    return eng.encrypt(v, level=level), v


@pytest.mark.parametrize("level", [0, 13, 31])
def test_hmult_chain_bit_exact(small, level):
    eng, ref, mode = small
    a, _ = _fresh(eng, level, seed=1)
    b, _ = _fresh(eng, level, seed=2)
    na, nb = eng.coefficient_domain_to_ntt_domain(a), eng.coefficient_domain_to_ntt_domain(b)
    ra = ref.coefficient_domain_to_ntt_domain(_np(a.data), level)
    rb = ref.coefficient_domain_to_ntt_domain(_np(b.data), level)
    q = np.array(eng.config.moduli, dtype=np.int64)[level: eng.config.num_q_primes].reshape(1, 1, -1, 1)
    if mode == "exact":
        np.testing.assert_array_equal(_np(na.data), ra)
    else:
        np.testing.assert_array_equal(_np(na.data), ra % q)
    prod = eng.multiply(na, nb)
    rprod = ref.multiply(ra, rb, level)
    np.testing.assert_array_equal(_np(prod.data), rprod)  # canonical products: mode-independent
    relin = eng.relinearize(prod)
    rrelin = ref.relinearize(rprod, _np(eng.relinearization_key.data), level)
    np.testing.assert_array_equal(_np(relin.data), rrelin)
    resc = eng.rescale_to_next_level(relin)
    np.testing.assert_array_equal(_np(resc.data), ref.rescale(rrelin, level))
    assert resc.level == level + 1


@pytest.mark.parametrize("step", [1, -7, 513])
def test_rotation_bit_exact(small, step):
    eng, ref, _ = small
    level = 3
    a, _ = _fresh(eng, level, seed=3)
    out = eng.rotate_by_step(a, step)
    key = eng.rotation_key(step)
    want = ref.rotate(_np(a.data), _np(key.data), key.galois_element, level)
    np.testing.assert_array_equal(_np(out.data), want)
    # Hoisted rotations gather ModUp'd digits in the NTT domain; their residues
    # differ from direct rotation by key-noise-level terms, exactly as in the
    # reference, so they are checked against the oracle's hoisted path.
    key2 = eng.rotation_key(2)
    hoisted = eng.rotate_many_with_keys(a, [key, key2])
    flat_c1 = _np(a.data)[1].reshape(-1, a.limb_count, eng.config.N)
    prepared = ref.prepare_rotation_digits(flat_c1, level)
    for got, k in zip(hoisted, (key, key2)):
        want_h = ref.rotate_hoisted(_np(a.data), prepared, _np(k.data), k.galois_element, level)
        np.testing.assert_array_equal(_np(got.data), want_h)
    # a single non-zero key falls back to the direct path
    single = eng.rotate_many_with_keys(a, [key])
    np.testing.assert_array_equal(_np(single[0].data), want)


def test_batched_rotation_per_entry_keys(small):
    """rotate_by_steps on a [2, 3, limb, N] batch equals rotating each entry alone."""
    eng, ref, _ = small
    level = 5
    a, _ = _fresh(eng, level, batch=(3,), seed=9)
    steps = [1, 17, -300]
    out = eng.rotate_by_steps(a, steps)
    for b, s in enumerate(steps):
        single = a.clone()
        single.data = a.data[:, b: b + 1].contiguous()
        want = eng.rotate_by_step(single, s)
        np.testing.assert_array_equal(_np(out.data[:, b: b + 1]), _np(want.data))


def test_state_errors(small):
    eng, _, _ = small
    a, _ = _fresh(eng, 0, batch=(1,))
    with pytest.raises(CkksStateError):
        eng.multiply(a, a)  # coefficient-domain inputs are rejected
    last = eng.config.public_level_count - 1
    b, _ = _fresh(eng, last, batch=(1,))
    with pytest.raises(MaximumLevelError):
        eng.rescale_to_next_level(b)


@pytest.fixture(scope="module")
def full():
    eng = CkksEngine(H.L34, ntt_mode="fast", rng_seed=11)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    return eng


def test_full_size_hmult_error(full):
    eng = full
    rng = np.random.default_rng(20260807)
    x = rng.uniform(-1, 1, eng.config.slots)
    y = rng.uniform(-1, 1, eng.config.slots)
    cx, cy = eng.encrypt(x), eng.encrypt(y)
    out = eng.hmult(cx, cy)
    got = eng.decrypt(out)
    err = np.abs(got.real - x * y).max()
    print(f"logN=16 L34 HMult max-abs error {err:.3e}")
    assert err < TOL_MULTIPLY


def test_full_size_rotation_error(full):
    eng = full
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, eng.config.slots)
    c = eng.encrypt(x, level=2)
    out = eng.rotate_by_step(c, 1)
    got = eng.decrypt(out)
    err = np.abs(got.real - np.roll(x, 1)).max()
    print(f"logN=16 L34 rotate(1) max-abs error {err:.3e}")
    assert err < TOL_KEYSWITCH


def test_basic_flow_l7():
    """README flow on preset slots8192-scale40-levels7 (logN=14): add, NTT,
    multiply, relinearize, rescale, rotate by 1, decrypt (atol 2e-5)."""

    eng = CkksEngine(H.L7, rng_seed=0x12345678)
    eng.generate_secret_key()
    x = np.linspace(-0.05, 0.05, 32)
    y = np.random.default_rng(20260807).uniform(-0.05, 0.05, 32)
    cx, cy = eng.encrypt(x), eng.encrypt(y)
    # addition of coefficient/standard ciphertexts (canonical add per component)
    csum = cx.clone()
    csum.data = torch.stack([eng.runtime.add_canonical(cx.data[i], cy.data[i], 0) for i in range(2)])
    s = eng.decrypt(csum)[:32].real
    np.testing.assert_allclose(s, x + y, atol=2e-5)
    prod = eng.rescale_to_next_level(eng.relinearize(eng.multiply(eng.coefficient_domain_to_ntt_domain(cx),
                                                                  eng.coefficient_domain_to_ntt_domain(cy))))
    p = eng.decrypt(prod)[:32].real
    np.testing.assert_allclose(p, x * y, atol=2e-5)
    rot = eng.rotate_by_step(prod, 1)
    r = eng.decrypt(rot).real
    np.testing.assert_allclose(r[1:33], (x * y), atol=2e-5)


@pytest.mark.parametrize("mode", ["fast", "exact"])
@pytest.mark.parametrize("level", [0, 7, 33])
def test_fused_keyswitch_matches_reference_sequence(mode, level):
    """fh_keyswitch_fused == the reference's per-digit op sequence, bit for bit,
    for a shared key and for per-entry keys, at logN=12 and logN=13."""
    for log_n in (12, 13):
        eng = CkksEngine(H.config(log_n), ntt_mode=mode, rng_seed=21)
        eng.generate_secret_key()
        key = eng.generate_relinearization_key()
        rng = np.random.default_rng(level)
        comp = H.to_dev(H.random_residues(rng, 3, H.moduli_array(eng.config, level, eng.config.num_q_primes - level),
                                          eng.config.N))
        base = H.to_dev(H.random_residues(rng, 3, H.moduli_array(eng.config, level, eng.config.num_q_primes - level),
                                          eng.config.N))
        fused = eng.key_switcher.key_switch_fused(comp, key, level, base0=base)
        c0, c1 = eng.key_switcher.apply_key_switch(comp, key, level)
        from fhelium_amd import kernels as K
        want0 = K.add_canonical(base, c0, eng.runtime.params(level))
        np.testing.assert_array_equal(_np(fused[0]), _np(want0))
        np.testing.assert_array_equal(_np(fused[1]), _np(c1))
        keys = [eng.rotation_key(s) for s in (1, 2, 3)]
        fused_multi = eng.key_switcher.key_switch_fused(comp, keys, level)
        m0, m1 = eng.key_switcher.apply_key_switch(comp, keys, level)
        np.testing.assert_array_equal(_np(fused_multi[0]), _np(m0))
        np.testing.assert_array_equal(_np(fused_multi[1]), _np(m1))


@pytest.fixture(scope="module")
def full_parity():
    eng = CkksEngine(H.config(16), rng_seed=17)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    return eng, EngineRef(H.host_tables(16))


@pytest.mark.parametrize("level", [0, 30])
def test_full_size_fused_hmult_and_rotation_bit_exact(full_parity, level):
    """logN=16, L34: the fused HMult (tensor product front, relinearization
    and rescale in the key switch) and a fused rotation reproduce the
    oracle's residues of the reference composition bit for bit."""

    eng, ref = full_parity
    a, _ = _fresh(eng, level, batch=(1,), seed=3)
    b, _ = _fresh(eng, level, batch=(1,), seed=4)
    got = eng.hmult(a, b)
    rprod = ref.multiply(ref.coefficient_domain_to_ntt_domain(_np(a.data), level),
                         ref.coefficient_domain_to_ntt_domain(_np(b.data), level), level)
    want = ref.rescale(ref.relinearize(rprod, _np(eng.relinearization_key.data), level), level)
    np.testing.assert_array_equal(_np(got.data), want)
    key = eng.rotation_key(5)
    rot = eng.rotate_by_step(got, 5)
    np.testing.assert_array_equal(_np(rot.data), ref.rotate(want, _np(key.data), key.galois_element, level + 1))


def test_rotation_composed_from_installed_keys():
    """Without a direct key and with key generation disallowed, rotate_by_step
    composes the rotation from installed keys along the reference's shortest
    walk (ckks_engine.py:3726-3751, core/rotation.py:57-110): residues equal
    the oracle's rotations applied hop by hop, and each hop adds at most the
    reference's key-switch tolerance."""
    cfg = H.config(12)
    eng = CkksEngine(cfg, rng_seed=11)
    eng.generate_secret_key()
    for s in (1, 4, -16, 64):
        eng.rotation_key(s)
    eng.allow_sk_gen = False
    ref = EngineRef(H.host_tables(12))
    level = 2
    a, v = _fresh(eng, level, batch=(3,), seed=21)
    assert eng.rotation_path(69) == [1, 4, 64]
    assert eng.rotation_path(-11) == [-16, 1, 4]
    assert eng.rotation_path(4) == [4]
    with pytest.raises(KeyError):
        eng.rotation_key(69)
    for step in (69, -11):
        out = eng.rotate_by_step(a, step)
        want = _np(a.data)
        for hop in eng.rotation_path(step):
            key = eng.rotation_keys[hop]
            want = ref.rotate(want, _np(key.data), key.galois_element, level)
        np.testing.assert_array_equal(_np(out.data), want, err_msg=f"step {step}")
        err = np.abs(eng.decrypt(out).real - np.roll(v, step, axis=-1)).max()
        assert err < 3 * TOL_KEYSWITCH, (step, err)
    # per-entry steps mixing direct and composed walks
    steps = [69, 4, -11]
    out = eng.rotate_by_steps(a, steps)
    for b, s in enumerate(steps):
        single = a.clone()
        single.data = a.data[:, b: b + 1].contiguous()
        np.testing.assert_array_equal(_np(out.data[:, b: b + 1]), _np(eng.rotate_by_step(single, s).data))
    # a step 0 entry is left as it is
    out0 = eng.rotate_by_steps(a, [0, 4, 69])
    np.testing.assert_array_equal(_np(out0.data[:, 0]), _np(a.data[:, 0]))


@pytest.mark.parametrize("level", [0, 6, 33])
def test_fused_hoisted_rotations_bit_exact(small, level):
    """Unbatched rotate_many_with_keys runs the fused hoisted key switch
    (fh_keyswitch_hoisted): every result equals the oracle's restatement of
    the reference's hoisted path (hybrid_keyswitch.py:373-504) bit for bit,
    and the batched rotate_hoisted holds the same residues."""
    eng, ref, _ = small
    a, v = _fresh(eng, level, batch=(), seed=30 + level)
    steps = [1, -7, 0, 513, 2, 1]
    keys = [eng.rotation_key(s) for s in steps if s != 0]
    got = eng.rotate_many_with_keys(a, keys)
    prepared = ref.prepare_rotation_digits(_np(a.data)[1].reshape(1, a.limb_count, eng.config.N), level)
    src = _np(a.data).reshape(2, 1, a.limb_count, eng.config.N)
    for out, k in zip(got, keys):
        want = ref.rotate_hoisted(src, prepared, _np(k.data), k.galois_element, level)
        np.testing.assert_array_equal(_np(out.data).reshape(want.shape), want, err_msg=f"step {k.rotation_step}")
    batch = eng.rotate_hoisted(a, steps)
    it = iter(got)
    for e, s in enumerate(steps):
        want = _np(a.data) if s == 0 else _np(next(it).data)
        np.testing.assert_array_equal(_np(batch.data[:, e]), want, err_msg=f"entry {e}")
    many = eng.rotate_many_by_steps(a, steps)
    for e, s in enumerate(steps):
        np.testing.assert_array_equal(_np(many[e].data), _np(batch.data[:, e]), err_msg=f"by steps {e}")
    err = np.abs(eng.decrypt(batch).real[3] - np.roll(v, 513)).max()
    assert err < TOL_KEYSWITCH, err


@pytest.fixture(scope="module")
def full_parity_exact():
    eng = CkksEngine(H.config(16), ntt_mode="exact", rng_seed=19)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    return eng, EngineRef(H.host_tables(16))


@pytest.mark.parametrize("level", [0, 21])
def test_full_size_exact_mode_bit_exact(full_parity_exact, level):
    """logN=16, L34, exact NTT arithmetic (the reference's lazy butterflies;
    the key switch's exact binary64 ModUp and MAC): NTT conversion equals
    the oracle's lazy values, and HMult (composition) and a rotation
    reproduce the oracle's residues bit for bit."""

    eng, ref = full_parity_exact
    print(f"exact mode logN=16 level={level}", flush=True)
    a, _ = _fresh(eng, level, batch=(1,), seed=13)
    b, _ = _fresh(eng, level, batch=(1,), seed=14)
    na = eng.coefficient_domain_to_ntt_domain(a)
    ra = ref.coefficient_domain_to_ntt_domain(_np(a.data), level)
    np.testing.assert_array_equal(_np(na.data), ra)
    got = eng.hmult(a, b)
    rprod = ref.multiply(ra, ref.coefficient_domain_to_ntt_domain(_np(b.data), level), level)
    want = ref.rescale(ref.relinearize(rprod, _np(eng.relinearization_key.data), level), level)
    np.testing.assert_array_equal(_np(got.data), want)
    key = eng.rotation_key(-7)
    rot = eng.rotate_by_step(got, -7)
    np.testing.assert_array_equal(_np(rot.data), ref.rotate(want, _np(key.data), key.galois_element, level + 1))

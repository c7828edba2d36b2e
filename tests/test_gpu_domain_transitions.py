"""Domain transitions as the reference's engine defines them
(fhelium/engine/ckks_engine.py:867-985, 2790-2835; its test
tests/test_ckks_operation_correctness.py:645-672): functional and in-place
forms give the same residues, in-place forms transform the payload in its
storage and update the state, plaintexts keep Montgomery residues in both
domains, and malformed sources are rejected before any launch."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.core import CkksStateError
from fhelium_amd.engine import CkksEngine
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["fast", "exact"])
def eng(request):
    e = CkksEngine(H.config(12), ntt_mode=request.param, rng_seed=31)
    e.generate_secret_key()
    return e


def _mod(eng, t, level):
    q = torch.tensor(H.moduli_array(eng.config, level, t.size(-2)), dtype=torch.int64, device=t.device)
    return torch.remainder(t, q.view(-1, 1))


@pytest.mark.parametrize("level", [0, 17])
def test_ciphertext_in_place_equals_functional(eng, level):
    x = np.random.default_rng(level).uniform(-1, 1, (2, eng.config.slots))
    ct = eng.encrypt(x, level=level)
    want = eng.coefficient_domain_to_ntt_domain(ct)
    assert ct.polynomial_domain == "coefficient" and ct.residue_representation == "standard"
    work = ct.clone()
    storage = work.data.data_ptr()
    got = eng.coefficient_domain_to_ntt_domain_(work)
    assert got is work and work.data.data_ptr() == storage
    assert (work.polynomial_domain, work.residue_representation) == ("ntt", "montgomery")
    assert torch.equal(work.data, want.data)
    back = eng.ntt_domain_to_coefficient_domain(want)
    again = eng.ntt_domain_to_coefficient_domain_(work)
    assert again is work and work.data.data_ptr() == storage
    assert (work.polynomial_domain, work.residue_representation) == ("coefficient", "standard")
    assert torch.equal(work.data, back.data)
    assert torch.equal(back.data, ct.data)  # canonical round trip


@pytest.mark.parametrize("level", [0, 9])
def test_plaintext_transitions(eng, level):
    x = np.random.default_rng(100 + level).uniform(-1, 1, (3, eng.config.slots))
    coef = eng.encode(x, level=level)
    ntt = eng.encode(x, level=level, polynomial_domain="ntt")
    got = eng.coefficient_domain_to_ntt_domain(coef)
    assert (got.polynomial_domain, got.residue_representation) == ("ntt", "montgomery")
    assert coef.polynomial_domain == "coefficient"
    assert torch.equal(_mod(eng, got.data, level), _mod(eng, ntt.data, level))
    work = coef.__class__(coef.data.clone(), coef.level, coef.scale)
    assert eng.coefficient_domain_to_ntt_domain_(work) is work
    assert torch.equal(work.data, got.data)
    back = eng.ntt_domain_to_coefficient_domain(got)
    assert (back.polynomial_domain, back.residue_representation) == ("coefficient", "montgomery")
    assert torch.equal(_mod(eng, back.data, level), _mod(eng, coef.data, level))
    assert eng.ntt_domain_to_coefficient_domain_(work) is work
    assert torch.equal(work.data, back.data)


def test_malformed_sources_are_rejected(eng):
    x = np.random.default_rng(7).uniform(-1, 1, eng.config.slots)
    source = eng.encrypt(x, level=0)
    bad_basis = source.clone()
    bad_basis.modulus_basis = "bad"
    with pytest.raises(ValueError, match="modulus_basis"):
        eng.coefficient_domain_to_ntt_domain_(bad_basis)
    ntt = eng.coefficient_domain_to_ntt_domain(source)
    bad_repr = ntt.clone()
    bad_repr.residue_representation = "standard"
    with pytest.raises(ValueError, match="Montgomery"):
        eng.ntt_domain_to_coefficient_domain_(bad_repr)
    with pytest.raises(CkksStateError):
        eng.coefficient_domain_to_ntt_domain(ntt)  # already in the NTT domain
    with pytest.raises(CkksStateError):
        eng.ntt_domain_to_coefficient_domain(source)  # already in the coefficient domain
    stale = source.clone()
    stale.level = 3  # four limbs too many for level 3
    with pytest.raises(ValueError, match="limbs"):
        eng.coefficient_domain_to_ntt_domain(stale)
    pt = eng.encode(x, level=0)
    pt.residue_representation = "standard"
    with pytest.raises(ValueError, match="Montgomery"):
        eng.coefficient_domain_to_ntt_domain(pt)
    with pytest.raises(TypeError):
        eng.coefficient_domain_to_ntt_domain(source.data)
    # nothing was transformed by the rejected calls
    assert torch.equal(bad_basis.data, source.data) and torch.equal(bad_repr.data, ntt.data)

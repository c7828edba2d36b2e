"""FakeTensor shapes of the registered operator schemas (CPU, no device):
the shape-changing ops describe their outputs explicitly, as the reference's
fake registrations do (fhelium/native/wrapper/fake_op.py, checked by
tests/test_native_operator_invariants.py:140-199), and every other
functional op returns its first operand's shape."""

from __future__ import annotations

import pytest
import torch
from torch._subclasses.fake_tensor import FakeTensorMode

from fhelium_amd import torch_ops  # noqa: F401  registers the schemas

RNS = torch.ops.fhelium_rns_ops
CKKS = torch.ops.fhelium_ckks_ops
NTT = torch.ops.fhelium_ntt_ops


@pytest.mark.parametrize("batch_shape", [(), (2,), (2, 3)])
def test_shape_changing_ops_have_fake_shapes(batch_shape):
    with FakeTensorMode():
        comps = torch.empty(*batch_shape, 3, 32, dtype=torch.int64, device="cuda")
        coeffs = torch.empty(2, 7, dtype=torch.int64, device="cuda")
        params = torch.empty(8, 7, dtype=torch.int64, device="cuda")
        extended = RNS.mixed_radix_basis_extend_to_montgomery(comps, coeffs, params, 7)
        assert extended.shape == (*batch_shape, 7, 32) and extended.device.type == "cuda"
        centered = torch.empty(*batch_shape, 32, dtype=torch.int64, device="cuda")
        twice = torch.empty(7, dtype=torch.int64, device="cuda")
        lifted = RNS.lift_centered_coefficients(centered, twice)
        assert lifted.shape == (*batch_shape, 7, 32)
        lhs = torch.empty(*batch_shape, 7, 32, dtype=torch.int64, device="cuda")
        compressed = torch.empty(*batch_shape, 7, 4, dtype=torch.int64, device="cuda")
        implicit = torch.empty(*batch_shape, 7, dtype=torch.int64, device="cuda")
        outs = (RNS.montgomery_mul_cyclic_compressed(lhs, compressed, params),
                RNS.montgomery_mul_contiguous_compressed(lhs, compressed, params),
                CKKS.add_cyclic_compressed_plaintext_component(lhs, compressed, params),
                CKKS.add_contiguous_compressed_plaintext_component(lhs, compressed, params),
                CKKS.add_strided_plaintext_component(lhs, compressed, implicit, params))
        assert all(o.shape == lhs.shape for o in outs)


def test_functional_and_inplace_fakes():
    with FakeTensorMode():
        x = torch.empty(2, 5, 64, dtype=torch.int64, device="cuda")
        params = torch.empty(8, 5, dtype=torch.int64, device="cuda")
        tw = torch.empty(5, 64, dtype=torch.int64, device="cuda")
        assert NTT.forward_ntt_to_montgomery_compact_grouped_smem(x, tw, params, 3).shape == x.shape
        outer = torch.empty(5, 63, dtype=torch.int64, device="cuda")
        roots = torch.empty(5, 4, dtype=torch.int64, device="cuda")
        assert NTT.forward_ntt_to_montgomery_power_of_two_radix_compact(x, outer, roots, params).shape == x.shape
        assert NTT.inverse_ntt_to_standard_compact_grouped_smem_(x, tw, params, 3) is None
        q = torch.empty(2, 3, 64, dtype=torch.int64, device="cuda")
        p = torch.empty(2, 2, 64, dtype=torch.int64, device="cuda")
        inv = torch.empty(2, 3, dtype=torch.int64, device="cuda")
        assert CKKS.keyswitch_moddown_qp_to_q(q, p, inv, params).shape == q.shape

"""The engine compositions of oracle/engine_ref.py on the reference's own CPU
kernels -- TEST AND BASELINE INFRASTRUCTURE ONLY.

``RefKernelEngine`` composes torch.ops.fhelium_{ntt,rns,ckks}_ops CPU kernels
(the reference's csrc/ops/*/cpu/*.cpp, built into oracle/_ref by
oracle/build_ref.py) in the order the reference's Python engine composes
them, with the same method names as ``EngineRef``:

  multiply            fhelium/engine/ckks_engine.py:1583-1598
  relinearize         ckks_engine.py:1651-1667 + hybrid_keyswitch.py:321-371
  key switch digit    hybrid_keyswitch.py:108-151, :254-298,
                      direct_keyswitch_consumer.py:55-82
  ModDown             hybrid_keyswitch.py:530-568
  rescale             ckks_rescale.py:339-377, 224-269
  rotate              ckks_engine.py:3822-3897, 3334-3350

The NTT is the reference's CPU backend: the indexed radix-2 schedules
(fhelium/engine/ntt/plans/indexed_radix2.py:58-126) with stage-expanded
twiddles (twiddles.py:87-106).  Load the library first
(``torch.ops.load_library``); it defines the same operator names as
fhelium_amd.torch_ops, so this module runs in a process of its own.
"""

from __future__ import annotations

import numpy as np
import torch

from fhelium_amd import rns

NTT = None
RNS = None
CKKS = None


def _ops():
    global NTT, RNS, CKKS
    NTT, RNS, CKKS = torch.ops.fhelium_ntt_ops, torch.ops.fhelium_rns_ops, torch.ops.fhelium_ckks_ops


def radix2_schedules(log_n: int):
    """Forward DIF and inverse DIT (even, odd, twiddle-index) rows of the
    reference's indexed radix-2 plan, as int32 tensors [stages, N/2]."""

    n = 1 << log_n
    half_n = n // 2
    f_even = np.empty((log_n, half_n), np.int32)
    f_tw = np.empty((log_n, half_n), np.int32)
    i_even = np.empty((log_n, half_n), np.int32)
    i_tw = np.empty((log_n, half_n), np.int32)
    f_half = np.empty(log_n, np.int64)
    i_half = np.empty(log_n, np.int64)
    for s in range(log_n):
        half = n >> (s + 1)
        g = np.repeat(np.arange(1 << s), half)
        k = np.tile(np.arange(half), 1 << s)
        f_even[s] = 2 * g * half + k
        f_tw[s] = (1 << s) + g
        f_half[s] = half
        half = 1 << s
        groups = n // (2 * half)
        g = np.repeat(np.arange(groups), half)
        k = np.tile(np.arange(half), groups)
        i_even[s] = 2 * g * half + k
        i_tw[s] = groups + g
        i_half[s] = half
    f_odd = (f_even + f_half[:, None]).astype(np.int32)
    i_odd = (i_even + i_half[:, None]).astype(np.int32)
    t = torch.from_numpy
    return t(f_even), t(f_odd), t(f_tw), t(i_even), t(i_odd), t(i_tw)


class RefKernelEngine:
    def __init__(self, tables: rns.HostTables) -> None:
        if NTT is None:
            _ops()
        self.t = tables
        self.cfg = tables.config
        self.params = tables.rns_params
        fwd, inv = rns.compact_twiddles_montgomery(self.cfg.moduli, self.cfg.logN)
        fe, fo, ft, ie, io, it = radix2_schedules(self.cfg.logN)
        self.sched_f = (fe, fo)
        self.sched_i = (ie, io)
        stages, bf = ft.shape
        self.fwd_tw = fwd.index_select(1, ft.reshape(-1).long()).reshape(fwd.size(0), stages, bf)
        self.inv_tw = inv.index_select(1, it.reshape(-1).long()).reshape(inv.size(0), stages, bf)

    def _rows(self, level: int, include_p: bool) -> slice:
        stop = self.cfg.total_num_primes if include_p else self.cfg.num_q_primes
        return slice(level, stop)

    def _p(self, level, include_p=False):
        return self.params[:, self._rows(level, include_p)].contiguous()

    @staticmethod
    def _flat(x):
        return x.reshape(-1, x.shape[-2], x.shape[-1])

    def ntt_forward(self, x, level, include_p=False, to_montgomery=True):
        v = self._flat(x).clone()
        tw = self.fwd_tw[self._rows(level, include_p)]
        op = NTT.forward_ntt_to_montgomery_indexed_ if to_montgomery else NTT.forward_ntt_montgomery_indexed_
        op(v, *self.sched_f, tw, self._p(level, include_p))
        return v.reshape(x.shape)

    def ntt_inverse_standard(self, x, level, include_p=False):
        v = self._flat(x).clone()
        NTT.inverse_ntt_to_standard_indexed_(v, *self.sched_i, self.inv_tw[self._rows(level, include_p)],
                                             self._p(level, include_p))
        return v.reshape(x.shape)

    def coefficient_domain_to_ntt_domain(self, c, level):
        return torch.stack([self.ntt_forward(comp, level) for comp in c])

    def multiply(self, a, b, level):
        p = self._p(level)
        f = self._flat
        sh = a.shape[1:]
        d0 = RNS.montgomery_mul(f(a[0]), f(b[0]), p)
        d1 = RNS.add_lazy(RNS.montgomery_mul(f(a[0]), f(b[1]), p), RNS.montgomery_mul(f(a[1]), f(b[0]), p), p)
        d2 = RNS.montgomery_mul(f(a[1]), f(b[1]), p)
        return torch.stack([d0.reshape(sh), d1.reshape(sh), d2.reshape(sh)])

    def key_switch(self, component, key, level):
        qp_p = self._p(level, True)
        dest = qp_p.size(1)
        acc0 = acc1 = None
        for spec in self.t.layout.digit_specs(level):
            rows = spec.component_row_ids
            src = component[:, rows[0]: rows[-1] + 1].contiguous()
            d = self.t.digit_tables[spec.prime_ids]
            digits = (RNS.mixed_radix_decompose(src, d.normalizers, d.propagation, d.modulus_lo, d.modulus_hi,
                                                d.neg_inv_lo, d.neg_inv_hi) if len(rows) > 1 else src)
            coef = d.basis_extension[:, level:].contiguous()
            if coef.numel() == 0:
                coef = torch.empty(0, dest, dtype=torch.int64)
            ext = RNS.mixed_radix_basis_extend_to_montgomery(digits, coef, qp_p, dest)
            ext = self.ntt_forward(ext, level, True, to_montgomery=False)
            if acc0 is None:
                acc0 = torch.zeros_like(ext)
                acc1 = torch.zeros_like(ext)
            CKKS.keyswitch_accumulate_digit_products_(acc0, acc1, ext, key[spec.key_digit_index], qp_p, level)
        return self.mod_down(acc0, acc1, level)

    def mod_down(self, acc0, acc1, level):
        p = self.cfg.num_p_primes
        qp_p = self._p(level, True)
        inv = self.t.moddown_inverses[level]
        out = []
        for acc in (acc0, acc1):
            std = self.ntt_inverse_standard(acc, level, True)
            out.append(CKKS.keyswitch_moddown_qp_to_q(std[:, :-p].contiguous(), std[:, -p:].contiguous(), inv, qp_p))
        return out[0], out[1]

    def relinearize(self, c3, key, level):
        b = c3.shape[1:]
        d = torch.stack([self.ntt_inverse_standard(x, level) for x in c3])
        f = self._flat
        corr0, corr1 = self.key_switch(f(d[2]), key, level)
        p = self._p(level)
        out0 = f(d[0]) + corr0
        out1 = f(d[1]) + corr1
        RNS.canonicalize_residues_(out0, p)
        RNS.canonicalize_residues_(out1, p)
        return torch.stack([out0.reshape(b), out1.reshape(b)])

    def rescale(self, c, level):
        ids = self.t.layout.prime_ids(level)
        inv = self.t.rescale_inverses[level][: len(ids) - 1].contiguous()
        sub = self.params[:, ids[1]: ids[-1] + 1].contiguous()
        half = self.cfg.moduli[ids[0]] // 2
        outs = []
        for comp in c:
            flat = self._flat(comp)
            r = CKKS.rescale_drop_leading_prime_nearest(flat[:, 1:].contiguous(), inv, flat[:, 0].contiguous(),
                                                         sub, half)
            outs.append(r.reshape(*comp.shape[:-2], comp.shape[-2] - 1, comp.shape[-1]))
        return torch.stack(outs)

    def rotate(self, c, key, galois_element, level):
        idx, sgn = rns.coefficient_galois_tables(self.cfg.N, galois_element)
        twice = 2 * torch.tensor(self.cfg.moduli[self._rows(level, False)], dtype=torch.int64)
        f = self._flat
        r0 = CKKS.apply_coefficient_galois_automorphism(f(c[0]).contiguous(), idx, sgn, twice)
        r1 = CKKS.apply_coefficient_galois_automorphism(f(c[1]).contiguous(), idx, sgn, twice)
        corr0, corr1 = self.key_switch(r1, key, level)
        out0 = RNS.add_canonical(r0, corr0, self._p(level))
        return torch.stack([out0.reshape(c.shape[1:]), corr1.reshape(c.shape[1:])])

"""CPU restatement of the reference engine methods -- TEST INFRASTRUCTURE ONLY.

Composes the oracle kernels (oracle.py) in exactly the order the reference's
Python engine composes its native ops, so that each GPU engine method can be
compared bit for bit with the reference CPU semantics:

  multiply            fhelium/engine/ckks_engine.py:1583-1598
  relinearize         ckks_engine.py:1651-1667 + hybrid_keyswitch.py:321-371
  key switch digit    hybrid_keyswitch.py:108-151 (mixed radix), :254-298
                      (basis extension), direct_keyswitch_consumer.py:55-82
                      (forward NTT + multiply-accumulate)
  ModDown             hybrid_keyswitch.py:530-568
  rescale             ckks_rescale.py:339-377, 224-269
  rotate              ckks_engine.py:3822-3897, 3334-3350
  hoisted rotations   hybrid_keyswitch.py:373-504

Scalar tables (rns_params, twiddles, ModUp / ModDown / rescale inverses) come
from ``fhelium_amd.rns.HostTables``; they are the reference's closed forms and
are pinned separately by tests/test_tables.py.  NTT here is the reference's
exact radix-2 schedule (lazy outputs), as in its CPU backend.
"""

from __future__ import annotations

import numpy as np

from fhelium_amd import rns
from oracle import oracle as O


class EngineRef:
    def __init__(self, tables: rns.HostTables) -> None:
        self.t = tables
        self.cfg = tables.config
        self.params = tables.rns_params.numpy()
        fwd, inv = rns.compact_twiddles_montgomery(self.cfg.moduli, self.cfg.logN)
        self.fwd = fwd.numpy()
        self.inv = inv.numpy()

    # -- row intervals ------------------------------------------------------
    def _rows(self, level: int, include_p: bool) -> slice:
        stop = self.cfg.total_num_primes if include_p else self.cfg.num_q_primes
        return slice(level, stop)

    def _p(self, level, include_p=False):
        return np.ascontiguousarray(self.params[:, self._rows(level, include_p)])

    def _tw(self, table, level, include_p=False):
        return np.ascontiguousarray(table[self._rows(level, include_p)])

    def ntt_forward(self, x, level, include_p=False, to_montgomery=True):
        sh = x.shape
        v = x.reshape(-1, sh[-2], sh[-1])
        out = O.ntt_forward(v, self._tw(self.fwd, level, include_p), self._p(level, include_p), to_montgomery)
        return out.reshape(sh)

    def ntt_inverse(self, x, level, include_p=False, out_form=2):
        sh = x.shape
        v = x.reshape(-1, sh[-2], sh[-1])
        out = O.ntt_inverse(v, self._tw(self.inv, level, include_p), self._p(level, include_p), out_form)
        return out.reshape(sh)

    # -- engine methods -------------------------------------------------------
    def coefficient_domain_to_ntt_domain(self, c, level):
        return np.stack([self.ntt_forward(comp, level) for comp in c])

    def multiply(self, a, b, level):
        p = self._p(level)
        f = lambda x: x.reshape(-1, x.shape[-2], x.shape[-1])
        sh = a.shape[1:]
        d0 = O.binary("mul", f(a[0]), f(b[0]), p)
        x0y1 = O.binary("mul", f(a[0]), f(b[1]), p)
        x1y0 = O.binary("mul", f(a[1]), f(b[0]), p)
        d1 = O.binary("add_lazy", x0y1, x1y0, p)
        d2 = O.binary("mul", f(a[1]), f(b[1]), p)
        return np.stack([d0.reshape(sh), d1.reshape(sh), d2.reshape(sh)])

    def key_switch(self, component, key, level):
        """component [B, Q_l, N] coefficient/standard -> (corr0, corr1)."""

        cfg = self.cfg
        qp_p = self._p(level, True)
        dest = qp_p.shape[1]
        acc0 = acc1 = None
        for spec in self.t.layout.digit_specs(level):
            rows = spec.component_row_ids
            src = np.ascontiguousarray(component[:, rows[0]: rows[-1] + 1])
            d = self.t.digit_tables[spec.prime_ids]
            if len(rows) > 1:
                digits = O.mixed_radix_decompose(src, d.normalizers.numpy(), d.propagation.numpy(),
                                                 d.modulus_lo.numpy(), d.modulus_hi.numpy(),
                                                 d.neg_inv_lo.numpy(), d.neg_inv_hi.numpy())
            else:
                digits = src
            coef = np.ascontiguousarray(d.basis_extension.numpy()[:, level:])
            ext = O.basis_extend(digits, coef, qp_p, dest)
            ext = self.ntt_forward(ext, level, True, to_montgomery=False)
            if acc0 is None:
                acc0 = np.zeros_like(ext)
                acc1 = np.zeros_like(ext)
            acc0, acc1 = O.keyswitch_mac(acc0, acc1, ext, key[spec.key_digit_index], qp_p, level)
        return self.mod_down(acc0, acc1, level)

    def mod_down(self, acc0, acc1, level):
        p = self.cfg.num_p_primes
        qp_p = self._p(level, True)
        inv = self.t.moddown_inverses[level].numpy()
        out = []
        for acc in (acc0, acc1):
            std = self.ntt_inverse(acc, level, True, out_form=2)
            out.append(O.moddown(std[:, :-p], std[:, -p:], inv, qp_p))
        return out[0], out[1]

    def relinearize(self, c3, key, level):
        b = c3.shape[1:]
        d = np.stack([self.ntt_inverse(x, level) for x in c3])
        flat = lambda x: x.reshape(-1, x.shape[-2], x.shape[-1])
        corr0, corr1 = self.key_switch(flat(d[2]), key, level)
        q = np.array(self.cfg.moduli, dtype=np.int64)[self._rows(level, False)]
        p = self._p(level)
        out0 = O.unary("canonicalize", flat(d[0]) + corr0, p)
        out1 = O.unary("canonicalize", flat(d[1]) + corr1, p)
        del q
        return np.stack([out0.reshape(b), out1.reshape(b)])

    def rescale(self, c, level):
        ids = self.t.layout.prime_ids(level)
        inv = self.t.rescale_inverses[level][: len(ids) - 1].numpy()
        sub = np.ascontiguousarray(self.params[:, ids[1]: ids[-1] + 1])
        half = self.cfg.moduli[ids[0]] // 2
        outs = []
        for comp in c:
            flat = comp.reshape(-1, comp.shape[-2], comp.shape[-1])
            r = O.rescale(flat[:, 1:], flat[:, 0], inv, sub, half, True)
            outs.append(r.reshape(*comp.shape[:-2], comp.shape[-2] - 1, comp.shape[-1]))
        return np.stack(outs)

    def prepare_rotation_digits(self, c1, level):
        """hybrid_keyswitch.py:373-417: ModUp + forward NTT of every digit."""

        qp_p = self._p(level, True)
        out = []
        for spec in self.t.layout.digit_specs(level):
            rows = spec.component_row_ids
            src = np.ascontiguousarray(c1[:, rows[0]: rows[-1] + 1])
            d = self.t.digit_tables[spec.prime_ids]
            if len(rows) > 1:
                digits = O.mixed_radix_decompose(src, d.normalizers.numpy(), d.propagation.numpy(),
                                                 d.modulus_lo.numpy(), d.modulus_hi.numpy(),
                                                 d.neg_inv_lo.numpy(), d.neg_inv_hi.numpy())
            else:
                digits = src
            coef = np.ascontiguousarray(d.basis_extension.numpy()[:, level:])
            ext = O.basis_extend(digits, coef, qp_p, qp_p.shape[1])
            out.append(self.ntt_forward(ext, level, True, to_montgomery=False))
        return out

    def rotate_hoisted(self, c, prepared, key, galois_element, level):
        """ckks_engine.py:3786-3820 with hybrid_keyswitch.py:419-504."""

        n = self.cfg.N
        qp_p = self._p(level, True)
        src = rns.ntt_galois_indices(self.cfg.logN, galois_element).numpy()
        acc0 = np.zeros_like(prepared[0])
        acc1 = np.zeros_like(prepared[0])
        for spec, digit in zip(self.t.layout.digit_specs(level), prepared):
            rotated = O.galois_ntt(digit, src)
            acc0, acc1 = O.keyswitch_mac(acc0, acc1, rotated, key[spec.key_digit_index], qp_p, level)
        corr0, corr1 = self.mod_down(acc0, acc1, level)
        idx, sgn = rns.coefficient_galois_tables(n, galois_element)
        twice = 2 * np.array(self.cfg.moduli, dtype=np.int64)[self._rows(level, False)]
        flat = lambda x: x.reshape(-1, x.shape[-2], x.shape[-1])
        r0 = O.galois_coefficient(flat(c[0]), idx.numpy(), sgn.numpy(), twice)
        out0 = O.binary("add_canonical", r0, corr0, self._p(level))
        return np.stack([out0.reshape(c.shape[1:]), corr1.reshape(c.shape[1:])])

    def rotate(self, c, key, galois_element, level):
        n = self.cfg.N
        idx, sgn = rns.coefficient_galois_tables(n, galois_element)
        twice = 2 * np.array(self.cfg.moduli, dtype=np.int64)[self._rows(level, False)]
        flat = lambda x: x.reshape(-1, x.shape[-2], x.shape[-1])
        r0 = O.galois_coefficient(flat(c[0]), idx.numpy(), sgn.numpy(), twice)
        r1 = O.galois_coefficient(flat(c[1]), idx.numpy(), sgn.numpy(), twice)
        corr0, corr1 = self.key_switch(r1, key, level)
        out0 = O.binary("add_canonical", r0, corr0, self._p(level))
        return np.stack([out0.reshape(c.shape[1:]), corr1.reshape(c.shape[1:])])

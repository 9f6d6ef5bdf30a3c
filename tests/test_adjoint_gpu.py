"""GPU checks of the hand-written adjoints against central finite differences
of the forward kernels (complex128).

Losses are chosen invariant under the simple update's singular-vector phases
(the gauge a network state does not see): the contracted two-site tensor,
the norms of the new bond's columns and the new lambdas.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import torch  # noqa: E402

from dequantizer_amd import _lib, local  # noqa: E402

DEV = "cuda"


def _cg(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


class _SuCase:
    """One simple update on a local handle: node 0 (controlling, bond slot i1)
    and node 1 (slot i2), one degree-1 stub per outer bond."""

    def __init__(self, d1, d2, i1, i2, chi, rank, dims1, dims2, seed, p=2):
        self.h = local._local_handle(d1, d2, i1, i2, chi, p)
        self.chi, self.p, self.r = chi, p, rank
        self.d = (d1, d2)
        self.slots = (i1, i2)
        rng = np.random.default_rng(seed)
        h = self.h
        T = np.zeros(h.tensor_elems, dtype=np.complex128)
        self.active = np.zeros(h.tensor_elems, dtype=bool)
        for node, dims in ((0, dims1), (1, dims2)):
            off = int(h.offsets[node])
            view = T[off: off + h.node_size(node)].reshape(*((chi,) * len(dims)), p)
            act = self.active[off: off + h.node_size(node)].reshape(view.shape)
            sl = tuple(slice(0, x) for x in dims)
            view[sl] = _cg(rng, tuple(dims) + (p,))
            act[sl] = True
        self.T = T
        lam = np.zeros((h.topo.n_edges, chi))
        self.lam_active = np.zeros_like(lam, dtype=bool)
        for node, dims in ((0, dims1), (1, dims2)):
            for e, dm in zip(h.slot_edges[node], dims):
                if not self.lam_active[e].any():
                    lam[e, :dm] = np.sort(rng.uniform(0.2, 1.0, dm))[::-1]
                    self.lam_active[e, :dm] = True
        self.lam = lam
        halves = np.zeros((1, 2, p ** 4), dtype=np.complex128)
        halves[0, 0, : rank * p * p] = _cg(rng, rank * p * p) / 2
        halves[0, 1, : rank * p * p] = _cg(rng, rank * p * p) / 2
        self.halves = halves
        self.meta = np.array([[0, 0, rank, 0]], dtype=np.int32)
        self.tape_bytes = int(_lib.lib().dq_su_tape_bytes(h.ptr, rank))

    def forward(self, T, lam, halves, tape=None):
        L = _lib.lib()
        dT = torch.from_numpy(T.copy()).to(DEV)
        dl = torch.from_numpy(lam.reshape(-1).copy()).to(DEV)
        dm = torch.from_numpy(self.meta).to(DEV)
        dh = torch.from_numpy(halves.copy()).to(DEV)
        err = torch.zeros(1, dtype=torch.float64, device=DEV)
        ent = torch.zeros(1, dtype=torch.float64, device=DEV)
        st = _lib.stream_handle()
        if tape is None:
            tape = torch.zeros(self.tape_bytes, dtype=torch.uint8, device=DEV)
        _lib.check(L.dq_simple_update_tape(self.h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(dl), _lib.ptr(dm), 1,
                                           _lib.ptr(dh), self.r, 1e-10, self.chi, _lib.ptr(err), _lib.ptr(ent),
                                           _lib.ptr(tape), st))
        return dT.cpu().numpy(), dl.cpu().numpy().reshape(lam.shape), tape

    def node(self, T, k):
        off = int(self.h.offsets[k])
        return T[off: off + self.h.node_size(k)].reshape(*((self.chi,) * self.d[k]), self.p)

    def loss_and_grads(self, Tout, lamout, W, wl, wc):
        """L = Re<W, Theta'> + sum wl lambda' + sum wc |column|^2 and its
        gradients with respect to the outputs."""
        A = np.moveaxis(self.node(Tout, 0), self.slots[0], -1)   # (others.., p, k)
        B = np.moveaxis(self.node(Tout, 1), self.slots[1], -1)
        lm = lamout[0]
        Af, Bf = A.reshape(-1, self.chi), B.reshape(-1, self.chi)
        theta = (Af * lm) @ Bf.T
        colsq = (np.abs(Af) ** 2).sum(axis=0)
        L = float(np.real(np.vdot(W, theta)) + wl @ lm + wc @ colsq)
        Abar = (W @ np.conj(Bf)) * lm + 2 * wc * Af
        Bbar = (W.T @ np.conj(Af)) * lm
        lbar = wl + np.real(np.einsum("xy,xj,yj->j", np.conj(W), Af, Bf))
        gT = np.zeros_like(Tout)
        for k, Xb in ((0, Abar), (1, Bbar)):
            g = np.moveaxis(Xb.reshape(np.moveaxis(self.node(Tout, k), self.slots[k], -1).shape), -1, self.slots[k])
            off = int(self.h.offsets[k])
            gT[off: off + self.h.node_size(k)] = g.reshape(-1)
        glam = np.zeros_like(lamout)
        glam[0] = lbar
        return L, gT, glam

    def vjp(self, T, lam, halves, tape, gT, glam):
        L = _lib.lib()
        dT = torch.from_numpy(T.copy()).to(DEV)
        dl = torch.from_numpy(lam.reshape(-1).copy()).to(DEV)
        dm = torch.from_numpy(self.meta).to(DEV)
        dh = torch.from_numpy(halves.copy()).to(DEV)
        gT_d = torch.from_numpy(gT.copy()).to(DEV)
        gl_d = torch.from_numpy(glam.reshape(-1).copy()).to(DEV)
        hb = torch.zeros_like(dh)
        _lib.check(L.dq_simple_update_vjp(self.h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(dl), _lib.ptr(dm), 1,
                                          _lib.ptr(dh), self.r, 1e-10, _lib.ptr(tape), _lib.ptr(gT_d), _lib.ptr(gl_d),
                                          _lib.ptr(hb), _lib.stream_handle()))
        return gT_d.cpu().numpy(), gl_d.cpu().numpy().reshape(lam.shape), hb.cpu().numpy()


CASES = [
    # d1, d2, i1, i2, chi, gate rank, dims node 0, dims node 1
    (3, 3, 1, 0, 4, 4, (4, 4, 4), (4, 4, 4)),
    (3, 2, 2, 1, 4, 2, (4, 3, 4), (2, 4)),
    (2, 3, 0, 2, 4, 2, (2, 3), (3, 3, 2)),
    (1, 3, 0, 1, 4, 2, (4,), (4, 4, 4)),
    (3, 3, 0, 2, 16, 2, (16, 16, 16), (16, 16, 16)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_simple_update_vjp_matches_central_differences(case):
    d1, d2, i1, i2, chi, rank, dims1, dims2 = CASES[case]
    c = _SuCase(d1, d2, i1, i2, chi, rank, dims1, dims2, seed=100 + case)
    rng = np.random.default_rng(7 + case)
    Tout, lamout, tape = c.forward(c.T, c.lam, c.halves)
    n_out = chi
    shapeA = np.moveaxis(c.node(Tout, 0), i1, -1).reshape(-1, chi).shape[0]
    shapeB = np.moveaxis(c.node(Tout, 1), i2, -1).reshape(-1, chi).shape[0]
    W = _cg(rng, (shapeA, shapeB)) / np.sqrt(shapeA * shapeB)
    wl = rng.standard_normal(n_out)
    wc = rng.standard_normal(n_out)
    L0, gT, glam = c.loss_and_grads(Tout, lamout, W, wl, wc)
    Tb, lb, hb = c.vjp(c.T, c.lam, c.halves, tape, gT, glam)

    def L_of(T, lam, halves):
        To, lo, _ = c.forward(T, lam, halves)
        return c.loss_and_grads(To, lo, W, wl, wc)[0]

    h = 1e-6
    for trial in range(3):
        dT = np.where(c.active, _cg(rng, c.T.shape), 0.0)
        dl = np.where(c.lam_active, rng.standard_normal(c.lam.shape), 0.0)
        dh = np.zeros_like(c.halves)
        dh[0, :, : rank * 4] = _cg(rng, (2, rank * 4))
        fd = (L_of(c.T + h * dT, c.lam + h * dl, c.halves + h * dh)
              - L_of(c.T - h * dT, c.lam - h * dl, c.halves - h * dh)) / (2 * h)
        an = (np.real(np.vdot(Tb[c.active], dT[c.active])) + float(np.sum(lb[c.lam_active] * dl[c.lam_active]))
              + np.real(np.vdot(hb, dh)))
        assert abs(fd - an) < 1e-5 * max(1.0, abs(fd)), (trial, fd, an)
        # each input separately
        for name, args, val in (("T", (c.T + h * dT, c.lam, c.halves, c.T - h * dT, c.lam, c.halves),
                                 np.real(np.vdot(Tb[c.active], dT[c.active]))),
                                ("lam", (c.T, c.lam + h * dl, c.halves, c.T, c.lam - h * dl, c.halves),
                                 float(np.sum(lb[c.lam_active] * dl[c.lam_active]))),
                                ("halves", (c.T, c.lam, c.halves + h * dh, c.T, c.lam, c.halves - h * dh),
                                 np.real(np.vdot(hb, dh)))):
            fd1 = (L_of(*args[:3]) - L_of(*args[3:])) / (2 * h)
            assert abs(fd1 - val) < 1e-5 * max(1.0, abs(fd1)), (name, trial, fd1, val)

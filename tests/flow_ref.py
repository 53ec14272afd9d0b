"""Plain PyTorch (CPU, float64) reference of the RQS coupling flow, with the
same parameter layout, masks and spline conventions as the HIP kernels.
Test infrastructure: used to check the GPU forward, inverse and gradients."""
import math

import numpy as np
import torch

EPS_W = 1e-3
EPS_D = 1e-3


def masks(L, D):
    out = []
    for l in range(L):
        pat = l % 4
        A, B = [], []
        for i in range(D):
            inA = (i % 2 == 0) if pat == 0 else (i % 2 == 1) if pat == 1 else (
                i < D // 2) if pat == 2 else (i >= D // 2)
            (A if inA else B).append(i)
        out.append((A, B))
    return out


def prm_stride(K):
    """Output columns per transformed dimension: the 3K - 1 spline parameters
    padded to a multiple of 8 (the padding columns are zero)."""
    return (3 * K - 1 + 7) // 8 * 8


def layout(D, L, K, H, depth):
    """Offsets of every parameter block in the flat vector, as the library lays them out."""
    ms = masks(L, D)
    maxdB = max(len(B) for _, B in ms)
    ldp = ((maxdB * prm_stride(K) + 63) // 64) * 64
    off = 0
    lay = []
    for A, B in ms:
        d = {"A": A, "B": B}
        d["W1"] = (off, (len(A), H))
        off += len(A) * H
        d["b1"] = (off, (H,))
        off += H
        d["Wh"] = []
        for _ in range(1, depth):
            d["Wh"].append(((off, (H, H)), (off + H * H, (H,))))
            off += H * H + H
        d["Wo"] = (off, (H, ldp))
        off += H * ldp
        d["bo"] = (off, (ldp,))
        off += ldp
        lay.append(d)
    return lay, off, ldp


def bf16(x):
    """Round to bfloat16 (round to nearest even) and back, as the BF16 MLP path
    stores its operands."""
    return x.to(torch.bfloat16).to(x.dtype)


class FlowRef:
    """bf16=True rounds what the BF16 MFMA path rounds: the hidden and output
    weights, and every hidden activation after its ReLU; sums stay in float64."""

    def __init__(self, theta, D, L, K, H, depth, bf16=False):
        self.D, self.L, self.K, self.H, self.depth = D, L, K, H, depth
        self.bf = bf16
        self.lay, self.n, self.ldp = layout(D, L, K, H, depth)
        assert len(theta) == self.n
        self.theta = torch.tensor(np.asarray(theta, dtype=np.float64), requires_grad=True)

    def _blk(self, spec):
        off, shape = spec
        n = int(np.prod(shape))
        return self.theta[off:off + n].view(*shape)

    def mlp(self, l, uA):
        d = self.lay[l]
        r = bf16 if self.bf else (lambda t: t)
        h = r(torch.relu(uA @ self._blk(d["W1"]) + self._blk(d["b1"])))
        for W, b in d["Wh"]:
            h = r(torch.relu(h @ r(self._blk(W)) + self._blk(b)))
        return h @ r(self._blk(d["Wo"])) + self._blk(d["bo"])

    def _knots(self, p):
        K = self.K
        w = EPS_W + (1 - K * EPS_W) * torch.softmax(p[..., :K], -1)
        h = EPS_W + (1 - K * EPS_W) * torch.softmax(p[..., K:2 * K], -1)
        dint = EPS_D + torch.nn.functional.softplus(p[..., 2 * K:3 * K - 1])
        one = torch.ones(p.shape[:-1] + (1,), dtype=p.dtype)
        d = torch.cat([one, dint, one], -1)
        X = torch.cat([torch.zeros_like(one), torch.cumsum(w, -1)[..., :-1]], -1)
        Y = torch.cat([torch.zeros_like(one), torch.cumsum(h, -1)[..., :-1]], -1)
        return w, h, d, X, Y

    def spline_fwd(self, p, x):
        """z, log dz/dx of the RQS at x (the bin found from the knots)."""
        w, h, d, X, Y = self._knots(p)
        K = self.K
        k = torch.clamp((x.unsqueeze(-1) >= (X + w)).sum(-1), max=K - 1)
        g = lambda t: t.gather(-1, k.unsqueeze(-1)).squeeze(-1)
        wk, hk, Xk, Yk, dk, dk1 = g(w), g(h), g(X), g(Y), g(d[..., :-1]), g(d[..., 1:])
        xi = torch.clamp((x - Xk) / wk, 0, 1)
        s = hk / wk
        t = xi * (1 - xi)
        Dn = s + (dk1 + dk - 2 * s) * t
        z = Yk + hk * (s * xi ** 2 + dk * t) / Dn
        Pn = dk1 * xi ** 2 + 2 * s * t + dk * (1 - xi) ** 2
        logJ = 2 * torch.log(s) + torch.log(Pn) - 2 * torch.log(Dn)
        return z, logJ

    def logq(self, u):
        """log q(u) in the density direction, u: (n, D) tensor."""
        cur = u
        tot = torch.zeros(u.shape[0], dtype=u.dtype)
        for l in range(self.L):
            A, B = self.lay[l]["A"], self.lay[l]["B"]
            prm = self.mlp(l, cur[:, A])
            nxt = cur.clone()
            for j, bi in enumerate(B):
                c0 = j * prm_stride(self.K)
                p = prm[:, c0:c0 + 3 * self.K - 1]
                z, lj = self.spline_fwd(p, cur[:, bi])
                nxt[:, bi] = z
                tot = tot + lj
            cur = nxt
        return tot, cur

    def grad(self, u, c):
        if self.theta.grad is not None:
            self.theta.grad.zero_()
        lq, _ = self.logq(torch.tensor(u, dtype=torch.float64))
        loss = -(torch.tensor(c, dtype=torch.float64) * lq).sum()
        loss.backward()
        return self.theta.grad.detach().numpy().copy(), lq.detach().numpy()

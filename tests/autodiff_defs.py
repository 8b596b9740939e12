"""The S, T and V tensors of the standard regime from their definitions, by torch autodiff (FP64).

Test infrastructure.  The reference states what its chain-rule code computes:

  S^x_b = d/dx_b <I>,  S^y_b = d/dy_b <I>                                  (chnr_dS.c:47-48)
  T^x_{a b d} = d/dx_b <(e_a e_d - delta_ad/4) I>,  T^y likewise with d/dy_b  (chnr_dT.c:182-183)
  V_{a b d} = d/dx_a (d/dx_b + d/dy_b) <e_d I>                              (chnr_dV.c:361)

with the averages written through the interpolated weight functions of (|x|, |y|, cos b):

  <I> = g0,
  <(e_a e_d - delta/4) I> = (x_a x_d - delta x^2/4) l1 + (y_a y_d - delta y^2/4) l2
                            + (x_a y_d + y_a x_d - delta x.y/2) l3,
     l3 = l4/(2 x^2 y^2) - y cb l2/x,
     l1 = 4/(3 x^4) (v1 - x^2 y^2 (cb^2 - 1/4) l2 - 3/2 x^3 y cb l3)        (chnr_dT.c:222-225)
  <e_d I> = x_d g1 + y_d g2,   g1 = g3 - (y cb/x) g2                        (chnr_dV.c:413-415)

In the swap regime (set_invariants, getff-new.c:341-350) the table is read at (|x|, |x-y|,
cos(x, x-y)) and the same forms hold with the vector y replaced by w = x - y; the reference's
XMYSWAP_* routines (chnr_dS.c:74-83, chnr_dT.c:146-180, chnr_dV.c:195-275) convert those
derivatives back to (x, y).  Here that conversion is left to autodiff of the definitions.

Weight functions are analytic stand-ins for the table: any smooth functions of (|x|, |y|, cb)
serve, since the chain rule holds for all of them.
"""
import numpy as np
import torch
from torch.func import grad, hessian, jacrev, vmap

torch.set_default_dtype(torch.float64)

# smooth, non-separable stand-ins for the six interpolated families
WFUN = {
    "g0": lambda x, y, c: torch.exp(-0.3 * x - 0.2 * y) * (1 + 0.4 * c + 0.3 * c * c * x * y) + 0.1 * torch.sin(x * y * c),
    "l2": lambda x, y, c: torch.exp(-0.25 * x * y) * (0.5 + 0.2 * c * x - 0.1 * c * c) + 0.05 * torch.cos(x + y * c),
    "l4": lambda x, y, c: (x * y) ** 2 * torch.exp(-0.2 * (x + y)) * (1 + 0.3 * c + 0.1 * c * c * y),
    "v1": lambda x, y, c: x ** 4 * torch.exp(-0.1 * x * y) * (0.7 - 0.2 * c + 0.15 * c * c) + 0.02 * x * y * c,
    "g2": lambda x, y, c: x * c * torch.exp(-0.2 * x - 0.3 * y) * (1 + 0.1 * y * c) + 0.03 * torch.sin(x + 0.5 * y + c),
    "g3": lambda x, y, c: torch.exp(-0.15 * x * y) * (0.4 + 0.3 * c * x + 0.2 * c * y * y) + 0.04 * torch.cos(x * c - y),
}


def table_args(xv, yv, swap):
    """(|x|, y', cb') at which the table is read: y' = |y| or |x - y| (swap)."""
    xv = np.asarray(xv, np.float64)
    w = xv - np.asarray(yv, np.float64) if swap else np.asarray(yv, np.float64)
    x, y = np.linalg.norm(xv), np.linalg.norm(w)
    return x, y, float(xv @ w) / (x * y)


def _inv(xv, yv, swap):
    w = xv - yv if swap else yv
    x = torch.sqrt(xv @ xv)
    y = torch.sqrt(w @ w)
    return x, y, (xv @ w) / (x * y), w


def rel(a, b):
    """Normwise relative difference of two blocks."""
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


BLOCKS = {"V": slice(0, 64), "Tx": slice(64, 128), "Ty": slice(128, 192), "Sx": slice(192, 196), "Sy": slice(196, 200)}


def block_errors(got, ref):
    return {k: rel(got[s], ref[s]) for k, s in BLOCKS.items()}


def stv_definition_batch(X, Y, swap, W=WFUN):
    """S/T/V of the rows of X, Y (one regime per call) from the definitions, in the oracle's layout:
    V[64], T^x[64], T^y[64], S^x[4], S^y[4], tensor index 16 a + 4 b + d."""
    X = torch.tensor(np.asarray(X, np.float64))
    Y = torch.tensor(np.asarray(Y, np.float64))
    eye = torch.eye(4)

    def G0(a, b):
        x, y, c, _ = _inv(a, b, swap)
        return W["g0"](x, y, c)

    def Tm(a, b):
        x, y, c, w = _inv(a, b, swap)
        l2, l4, v1 = W["l2"](x, y, c), W["l4"](x, y, c), W["v1"](x, y, c)
        l3 = l4 / (2 * x * x * y * y) - y * c * l2 / x
        l1 = 4 / (3 * x ** 4) * (v1 - x * x * y * y * (c * c - 0.25) * l2 - 1.5 * x ** 3 * y * c * l3)
        t3 = torch.outer(a, a) - eye * (a @ a) / 4
        t4 = torch.outer(w, w) - eye * (w @ w) / 4
        t5 = torch.outer(a, w) + torch.outer(w, a) - 0.5 * eye * (a @ w)
        return t3 * l1 + t4 * l2 + t5 * l3

    def E(a, b):
        x, y, c, w = _inv(a, b, swap)
        g2, g3 = W["g2"](x, y, c), W["g3"](x, y, c)
        return a * (g3 - y * c / x * g2) + w * g2

    def F(a, b):
        return jacrev(E, argnums=0)(a, b) + jacrev(E, argnums=1)(a, b)

    def one(a, b):
        Sx = grad(G0, argnums=0)(a, b)
        Sy = grad(G0, argnums=1)(a, b)
        Tx = jacrev(Tm, argnums=0)(a, b).permute(0, 2, 1)
        Ty = jacrev(Tm, argnums=1)(a, b).permute(0, 2, 1)
        V = jacrev(F, argnums=0)(a, b).permute(2, 1, 0)
        return torch.cat([V.reshape(-1), Tx.reshape(-1), Ty.reshape(-1), Sx, Sy])

    return vmap(one)(X, Y).numpy()


def inject_batch(X, Y, swap, W=WFUN):
    """The 34 interpolated weight functions (order of oq_stv_inject) of the rows of X, Y (one regime
    per call): value, gradient and Hessian of the stand-ins at the table arguments."""
    args = np.array([table_args(a, b, swap) for a, b in zip(X, Y)])
    P = torch.tensor(args)
    out = []
    for k in ("g0", "l2", "l4", "v1", "g2", "g3"):
        f = W[k]
        g3 = lambda q, f=f: f(q[0], q[1], q[2])
        v = vmap(g3)(P)
        g = vmap(grad(g3))(P)
        h = vmap(hessian(g3))(P)
        out.append(dict(v=v, x=g[:, 0], y=g[:, 1], c=g[:, 2], xx=h[:, 0, 0], xy=h[:, 0, 1], xc=h[:, 0, 2],
                        yy=h[:, 1, 1], yc=h[:, 1, 2], cc=h[:, 2, 2]))
    g0, l2, l4, v1, g2, g3 = out
    z = torch.zeros(len(X))
    cols = [g0["c"], g0["x"], g0["y"], l2["x"], l2["y"], l2["v"], l2["c"], l4["x"], l4["y"], l4["v"], l4["c"],
            v1["x"], v1["y"], v1["v"], v1["c"], g2["v"], g2["y"], g2["c"], g2["yc"], g2["cc"], g2["xx"], g2["x"],
            g2["xy"], g2["xc"], g3["xx"], g3["x"], g3["xy"], g3["xc"], g3["c"], g3["yc"], g3["cc"], g3["y"],
            g2["yy"] if swap else z, g3["yy"] if swap else z]
    return torch.stack(cols, dim=1).numpy()

"""CPU restatement of the warm-start graph network (test infrastructure only).

The network has no counterpart in the reference; it is specified by
include/m3gnn.h.  Two modes:

* ``exact=False``: plain float64 arithmetic, no BF16 storage.  Pinned by
  central finite differences of the loss (tests/test_oracle_gnn.py), which
  fixes the backward formulas independently of how they are written here.
* ``exact=True``: the device's arithmetic, step by step: FP32 values, every
  product-accumulate a fused multiply-add in the device's order (bias
  first, then input index ascending), BF16 round-to-nearest-even storage of
  the activations, neighbour sums in adjacency order.  The fused
  multiply-add is evaluated as float64 product plus float64 sum rounded once
  to float32 (the float64 product of two float32 values is exact).  Weight
  gradients are returned as float64 sums.

Parameter layout: We (8x64), be, per layer W (128x64), b, Wd (64x4), bd.
"""
import numpy as np

NF, W, NO, NL = 8, 64, 4, 3
LSZ = 2 * W * W + W
OFF_WE, OFF_BE = 0, NF * W
OFF_L = OFF_BE + W
OFF_WD = OFF_L + NL * LSZ
OFF_BD = OFF_WD + W * NO
NPARAM = OFF_BD + NO
f32, f64 = np.float32, np.float64


def bf16_round(x):
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=f32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    s = (u >> 16) | np.where((u & 0xFFFF) != 0, np.uint64(0x40), np.uint64(0))
    out = np.where(special, s, r).astype(np.uint32) << 16
    return out.view(f32)


def fma32(a, b, c):
    return (a.astype(f64) * b.astype(f64) + c.astype(f64)).astype(f32)


def unpack(p):
    p = np.asarray(p)
    We = p[OFF_WE:OFF_BE].reshape(NF, W)
    be = p[OFF_BE:OFF_L]
    layers = []
    for l in range(NL):
        o = OFF_L + l * LSZ
        layers.append((p[o:o + 2 * W * W].reshape(2 * W, W), p[o + 2 * W * W:o + LSZ]))
    Wd = p[OFF_WD:OFF_BD].reshape(W, NO)
    bd = p[OFF_BD:OFF_BD + NO]
    return We, be, layers, Wd, bd


def _neighbour_sum(ptr, col, V, dtype):
    """sum over the neighbours of each vertex, in adjacency order"""
    n = len(ptr) - 1
    deg = np.diff(ptr)
    s = np.zeros((n, V.shape[1]), dtype=dtype)
    for k in range(int(deg.max())):
        rows = np.nonzero(deg > k)[0]
        s[rows] = (s[rows] + V[col[ptr[rows] + k]]).astype(dtype)
    return s


def _dense(A, Wm, b, exact):
    if not exact:
        return A @ Wm + b
    acc = np.broadcast_to(b.astype(f32), (A.shape[0], Wm.shape[1])).copy()
    for k in range(A.shape[1]):
        acc = fma32(A[:, k:k + 1], Wm[k][None, :], acc)
    return acc


def forward(X, params, norm, ptr, col, exact=True):
    """Returns y and the stored activations."""
    dt = f32 if exact else f64
    X = np.asarray(X, dtype=dt)
    params = np.asarray(params, dtype=dt)
    norm = np.asarray(norm, dtype=dt)
    We, be, layers, Wd, bd = unpack(params)
    mu, sig = norm[:NF], norm[NF:2 * NF]
    rnd = bf16_round if exact else (lambda v: v)
    Xn = ((X - mu) / sig).astype(dt)
    H = [rnd(np.maximum(_dense(Xn, We, be, exact), 0))]
    deg = np.diff(ptr).astype(dt)[:, None]
    Ms, Zs = [None], [None]
    for l in range(NL):
        Wl, bl = layers[l]
        M = rnd((_neighbour_sum(ptr, col, H[-1], dt) / deg).astype(dt))
        pre = _dense(np.concatenate([H[-1], M], axis=1), Wl, bl, exact)
        z = np.maximum(pre, 0).astype(dt)
        Zs.append(rnd(z))
        H.append(rnd((H[-1] + z).astype(dt)))
        Ms.append(M)
    Y = _dense(H[-1], Wd, bd, exact)
    return Y, dict(Xn=Xn, H=H, M=Ms, Z=Zs)


def loss_and_grad(X, Yt, params, norm, ptr, col, exact=True):
    """MSE loss of y against Yt / out_sigma, its gradient (float64 array in
    the parameter layout) and the prediction."""
    dt = f32 if exact else f64
    Y, a = forward(X, params, norm, ptr, col, exact)
    norm_d = np.asarray(norm, dtype=dt)
    params_d = np.asarray(params, dtype=dt)
    We, be, layers, Wd, bd = unpack(params_d)
    n = X.shape[0]
    ne = n * NO
    d = (Y - (np.asarray(Yt, dtype=dt) / norm_d[2 * NF:])).astype(dt)
    loss = float(np.sum(d.astype(f64) ** 2) / ne)
    scale = dt(2.0 / ne)
    dY = (d * scale).astype(dt)
    g = np.zeros(NPARAM)

    def wgrad(A, D, off):
        A64, D64 = A.astype(f64), D.astype(f64)
        K, N = A.shape[1], D.shape[1]
        g[off:off + K * N] = (A64.T @ D64).ravel()
        g[off + K * N:off + (K + 1) * N] = D64.sum(axis=0)

    def tdense(D, Wt):  # D (n x N) times Wt (N x K) in the device's order
        if not exact:
            return D @ Wt
        acc = np.zeros((D.shape[0], Wt.shape[1]), dtype=f32)
        for c in range(D.shape[1]):
            acc = fma32(D[:, c:c + 1], Wt[c][None, :], acc)
        return acc

    wgrad(a["H"][NL], dY, OFF_WD)
    dH = tdense(dY, Wd.T)
    deg = np.diff(ptr).astype(dt)[:, None]
    for l in range(NL, 0, -1):
        Wl, _ = layers[l - 1]
        dP = np.where(a["Z"][l] > 0, dH, 0).astype(dt)
        wgrad(np.concatenate([a["H"][l - 1], a["M"][l]], axis=1), dP, OFF_L + (l - 1) * LSZ)
        sa = tdense(dP, Wl[:W].T)
        sb = tdense(dP, Wl[W:].T)
        dHn = (dH + sa).astype(dt)
        dHn = (dHn + _neighbour_sum(ptr, col, (sb / deg).astype(dt), dt)).astype(dt)
        dH = dHn
    dP0 = np.where(a["H"][0] > 0, dH, 0).astype(dt)
    wgrad(a["Xn"], dP0, OFF_WE)
    return loss, g, Y


def features(T, T_prev, Visc, U_prev, P_prev, rn):
    """Per-vertex features and the state scales (s_u, s_p, s_T) of
    include/m3gnn.h."""
    n = len(T)
    Uv = np.asarray(U_prev[:3 * n]).reshape(n, 3)
    dT = np.asarray(T) - np.asarray(T_prev)
    su = np.sqrt(np.sum(Uv ** 2) / n)
    sp = np.sqrt(np.sum(np.asarray(P_prev) ** 2) / n)
    st = np.sqrt(np.sum(dT ** 2) / n)
    su = su if su > 0 else 1.0
    sp = sp if sp > 0 else 1.0
    st = st if st > 0 else 1.0
    X = np.zeros((n, NF), dtype=f32)
    X[:, 0] = T
    X[:, 1] = np.log(Visc[:n])
    X[:, 2:5] = Uv / su
    X[:, 5] = np.asarray(P_prev) / sp
    X[:, 6] = rn
    X[:, 7] = dT / st
    return X, np.array([su, sp, st])


def warm_start(Y, norm, scales, U_prev, P_prev, parent):
    """U0, P0 from the normalised prediction Y (n x 4)."""
    n = Y.shape[0]
    osig = np.asarray(norm, dtype=f64)[2 * NF:]
    dU = Y[:, :3].astype(f64) * osig[:3] * (scales[0] * scales[2])
    U0 = np.array(U_prev, dtype=f64)
    U0[:3 * n] += dU.ravel()
    U0[3 * n:] += (0.5 * (dU[parent[:, 0]] + dU[parent[:, 1]])).ravel()
    P0 = np.asarray(P_prev, dtype=f64) + Y[:, 3].astype(f64) * osig[3] * (scales[1] * scales[2])
    return U0, P0


def extrapolate(times, states, t, order):
    """Base state of the history entry points (include/m3gnn.h): the
    polynomial in time of degree min(order, len(times) - 1) through the most
    recent states (``times`` increasing, ``states`` in the same order),
    evaluated at ``t`` in Lagrange form; summed from the most recent state
    back, as the device does."""
    k = min(order + 1, len(times))
    ts = [float(x) for x in times[::-1][:k]]
    xs = [np.asarray(x, dtype=f64) for x in states[::-1][:k]]
    w = []
    for j in range(k):
        wj = 1.0
        for m in range(k):
            if m != j:
                wj *= (t - ts[m]) / (ts[j] - ts[m])
        w.append(wj)
    out = w[0] * xs[0]
    for j in range(1, k):
        out = out + w[j] * xs[j]
    return out

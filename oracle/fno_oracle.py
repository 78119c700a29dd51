"""CPU oracle of the spectral neural-operator field surrogate (test infrastructure).

Only the tests, the smoke entry and bench.py's CPU leg use this module, and
only as the checker of the HIP path in csrc/mcpm_fno.hip.  It restates, in
float64 numpy, the network the library trains and evaluates, with the same
BF16 roundings at the same places, so that the GPU forward pass, loss and
gradients can be compared with it within FP32-accumulation tolerance.

The network stands in for the periodic plane's screened-Poisson solve
(compute_SP_2D, FEM.cpp:249-610): nodal density in, the Q1 FEM solution of
(K + lambda^2 M) phi = M n (before the core subtraction of FEM.cpp:497-554)
out.  The reference has no such network; its definition is this module's:

  u        = (n - in_shift) * in_scale              (unique nodes, p = j N0 + i)
  h0[c]    = bf16(wl[c] u + bl[c])                  lift, width C = 64
  layer l  : A   = Tf . h_l            (BF16 x BF16 -> FP32; DFT along x)
             X   = DFT_y(A)            (FP32; 2m retained ky, m retained kx)
             Y   = X R_l               (complex 64 x 64 per mode, FP32)
             Cp  = bf16(IDFT_y(Y) * s) (s = c_kx / (N0 N1), c_0 = 1, else 2)
             z_l = bf16(W_l) . h_l + Tinv . Cp + b_l     (FP32)
             h_{l+1} = bf16(gelu(z_l))  (l < 3)
  y        = wq . z_3 + bq
  phi      = y / out_scale + out_shift

Retained modes: kx in [0, m); ky(a) = a for a < m, N1 - 2m + a for a >= m.
Tf[k'][x] = cos(2 pi k' x / N0) for k' < m, -sin(2 pi (k'-m) x / N0) for
k' >= m, rounded to BF16; Tinv = Tf transposed.  gelu is the tanh form.

Loss (normalised units, e = y - t, periodic forward differences):
  L = sum_p (|grad e|^2 + w0 e^2) / sum_p (|grad t|^2 + w0 t^2),
so it weights the error the way the force (the gradient of the
interpolated field) sees it.

exact=True skips every BF16 rounding; with it the backward pass is checked
against central finite differences, and the spectral path against numpy's
FFT (both in tests/test_fno_oracle.py).
"""
import numpy as np

C = 64
LAYERS = 4
GELU_K = np.sqrt(2.0 / np.pi)


def bf16(x, exact=False):
    """Round to BF16 (round to nearest even, via float32) and widen."""
    if exact:
        return np.asarray(x, dtype=np.float64)
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def layout(m):
    """Offsets of the flat FP32 parameter vector (shared with the library)."""
    off = {"wl": 0, "bl": C}
    base = 2 * C
    rsz = 2 * m * m * C * C * 2
    for l in range(LAYERS):
        off[f"W{l}"] = base
        off[f"b{l}"] = base + C * C
        off[f"R{l}"] = base + C * C + C
        base += C * C + C + rsz
    off["wq"] = base
    off["bq"] = base + C
    n = base + C + 1
    off["n"] = (n + 3) // 4 * 4
    return off


def unpack(params, m):
    off = layout(m)
    P = {"wl": params[off["wl"]:off["wl"] + C], "bl": params[off["bl"]:off["bl"] + C],
         "wq": params[off["wq"]:off["wq"] + C], "bq": params[off["bq"]]}
    for l in range(LAYERS):
        P[f"W{l}"] = params[off[f"W{l}"]:off[f"W{l}"] + C * C].reshape(C, C)  # [o][i]
        P[f"b{l}"] = params[off[f"b{l}"]:off[f"b{l}"] + C]
        r = params[off[f"R{l}"]:off[f"R{l}"] + 2 * m * m * C * C * 2].reshape(2 * m, m, C, C, 2)
        P[f"R{l}"] = r[..., 0] + 1j * r[..., 1]  # [a][kx][i][o]
    return P


def init_params(m, seed=0):
    """Initial weights (numpy PCG64): lift and projection U(-1,1), W and b
    U(-1/8, 1/8) (1/sqrt(C)), spectral weights U(-1,1)/C^2 per component."""
    rng = np.random.default_rng(seed)
    off = layout(m)
    p = np.zeros(off["n"], dtype=np.float32)
    p[off["wl"]:off["wl"] + C] = rng.uniform(-1, 1, C)
    p[off["bl"]:off["bl"] + C] = rng.uniform(-1, 1, C)
    a = 1.0 / np.sqrt(C)
    for l in range(LAYERS):
        p[off[f"W{l}"]:off[f"W{l}"] + C * C] = rng.uniform(-a, a, C * C)
        p[off[f"b{l}"]:off[f"b{l}"] + C] = rng.uniform(-a, a, C)
        n = 2 * m * m * C * C * 2
        p[off[f"R{l}"]:off[f"R{l}"] + n] = rng.uniform(-1, 1, n) / (C * C)
    p[off["wq"]:off["wq"] + C] = rng.uniform(-a, a, C)
    p[off["bq"]] = 0.0
    return p


def ky_of(a, m, n1):
    return a if a < m else n1 - 2 * m + a


def tables(n0, n1, m, exact=False):
    """Tf[k'][x] (BF16 values), and cos/sin of 2 pi ky(a) y / N1 (FP32)."""
    x = np.arange(n0)
    tf = np.zeros((2 * m, n0))
    for k in range(m):
        r = (k * x) % n0
        tf[k] = np.cos(2 * np.pi * r / n0).astype(np.float32)
        tf[m + k] = (-np.sin(2 * np.pi * r / n0)).astype(np.float32)
    tf = bf16(tf, exact)
    y = np.arange(n1)
    cy = np.zeros((2 * m, n1))
    sy = np.zeros((2 * m, n1))
    for a in range(2 * m):
        r = (ky_of(a, m, n1) * y) % n1
        cy[a] = np.cos(2 * np.pi * r / n1).astype(np.float32)
        sy[a] = np.sin(2 * np.pi * r / n1).astype(np.float32)
    return tf, cy, sy


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_K * (z + 0.044715 * z ** 3)))


def gelu_d(z):
    t = np.tanh(GELU_K * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_K * (1.0 + 3 * 0.044715 * z * z)


def kx_scale(m, n0, n1, adjoint_inverse):
    s = np.ones(m)
    if adjoint_inverse:
        s[:] = 2.0 / (n0 * n1)
        s[0] = 1.0 / (n0 * n1)
    return s


def dft_x(h, tf):
    """h[y][x][c] -> A[y][k'][c] = sum_x Tf[k'][x] h[y][x][c]."""
    return np.einsum("kx,yxc->ykc", tf, h)


def dft_y(A, cy, sy, m, s):
    """X[a][kx][c] = s[kx] sum_y A[y][kx] e^{-i phi(a,y)}, A complex from packed."""
    are, aim = A[:, :m, :], A[:, m:, :]
    xr = np.einsum("ay,ykc->akc", cy, are) + np.einsum("ay,ykc->akc", sy, aim)
    xi = np.einsum("ay,ykc->akc", cy, aim) - np.einsum("ay,ykc->akc", sy, are)
    return (xr + 1j * xi) * s[None, :, None]


def idft_y(Y, cy, sy, m, s, exact):
    """Cp[y][k'][o]: s[kx] sum_a Y[a][kx] e^{+i phi(a,y)}, packed re | im, BF16."""
    cr = np.einsum("ay,ako->yko", cy, Y.real) - np.einsum("ay,ako->yko", sy, Y.imag)
    ci = np.einsum("ay,ako->yko", cy, Y.imag) + np.einsum("ay,ako->yko", sy, Y.real)
    cr = cr * s[None, :, None]
    ci = ci * s[None, :, None]
    return bf16(np.concatenate([cr, ci], axis=1), exact)


def combine(h, W, tf, cp, b):
    """z[y][x][o] = sum_i W[o][i] h[y][x][i] + sum_k' Tf[k'][x] Cp[y][k'][o] + b[o]."""
    return np.einsum("oi,yxi->yxo", W, h) + np.einsum("kx,yko->yxo", tf, cp) + b


def forward(params, u, m, exact=False):
    """u: [N1][N0] normalised density.  Returns y [N1][N0] and the cache."""
    n1, n0 = u.shape
    P = unpack(np.asarray(params, dtype=np.float64), m)
    tf, cy, sy = tables(n0, n1, m, exact)
    s_fwd = np.ones(m)
    s_inv = kx_scale(m, n0, n1, True)
    h = bf16(P["wl"][None, None, :] * u[:, :, None] + P["bl"], exact)
    cache = {"u": u, "h": [h], "z": [], "X": [], "tf": tf, "cy": cy, "sy": sy, "P": P}
    for l in range(LAYERS):
        A = dft_x(h, tf)
        X = dft_y(A, cy, sy, m, s_fwd)
        Y = np.einsum("aki,akio->ako", X, P[f"R{l}"])
        cp = idft_y(Y, cy, sy, m, s_inv, exact)
        z = combine(h, bf16(P[f"W{l}"], exact), tf, cp, P[f"b{l}"])
        cache["X"].append(X)
        cache["z"].append(z)
        if l < LAYERS - 1:
            h = bf16(gelu(z), exact)
            cache["h"].append(h)
    y = z @ P["wq"] + P["bq"]
    return y, cache


def grad2(e):
    dx = np.roll(e, -1, axis=1) - e
    dy = np.roll(e, -1, axis=0) - e
    return dx * dx + dy * dy


def loss_denominator(t, w0):
    return float(np.sum(grad2(t) + w0 * t * t))


def loss_and_dy(y, t, w0):
    e = y - t
    D = loss_denominator(t, w0)
    L = float(np.sum(grad2(e) + w0 * e * e)) / D
    lap = 4 * e - np.roll(e, 1, 0) - np.roll(e, -1, 0) - np.roll(e, 1, 1) - np.roll(e, -1, 1)
    return L, 2.0 * (lap + w0 * e) / D


def backward(params, cache, dy, m, exact=False):
    """Gradient of the flat parameter vector given dL/dy."""
    P = cache["P"]
    u = cache["u"]
    n1, n0 = u.shape
    tf, cy, sy = cache["tf"], cache["cy"], cache["sy"]
    s_one = np.ones(m)
    s_adj = kx_scale(m, n0, n1, True)
    off = layout(m)
    g = np.zeros(off["n"])
    z3 = cache["z"][LAYERS - 1]
    g[off["wq"]:off["wq"] + C] = np.einsum("yx,yxc->c", dy, z3)
    g[off["bq"]] = dy.sum()
    dz = dy[:, :, None] * P["wq"][None, None, :]
    for l in range(LAYERS - 1, -1, -1):
        h = cache["h"][l]
        dzb = bf16(dz, exact)
        g[off[f"W{l}"]:off[f"W{l}"] + C * C] = np.einsum("yxo,yxi->oi", dzb, h).ravel()
        g[off[f"b{l}"]:off[f"b{l}"] + C] = dz.sum(axis=(0, 1))
        dC = dft_x(dzb, tf)                       # adjoint of the x inverse
        dY = dft_y(dC, cy, sy, m, s_adj)          # adjoint of the y inverse
        X = cache["X"][l]
        dR = np.conj(X)[:, :, :, None] * dY[:, :, None, :]
        r = np.stack([dR.real, dR.imag], axis=-1).ravel()
        g[off[f"R{l}"]:off[f"R{l}"] + r.size] = r
        dX = np.einsum("ako,akio->aki", dY, np.conj(P[f"R{l}"]))
        dA = idft_y(dX, cy, sy, m, s_one, exact)  # adjoint of the forward y DFT
        dh = np.einsum("oi,yxo->yxi", bf16(P[f"W{l}"], exact), dzb) + np.einsum("kx,yki->yxi", tf, dA)
        if l > 0:
            dz = dh * gelu_d(cache["z"][l - 1])
        else:
            g[off["wl"]:off["wl"] + C] = np.einsum("yxc,yx->c", dh, u)
            g[off["bl"]:off["bl"] + C] = dh.sum(axis=(0, 1))
    return g


def loss_grad(params, u, t, m, w0, exact=False):
    y, cache = forward(params, u, m, exact)
    L, dy = loss_and_dy(y, t, w0)
    return L, backward(params, cache, dy, m, exact), y


# ---- the field the network learns: the periodic Q1 FEM solve in closed form ----

def fem_transfer(n0, n1, h, lam):
    """Transfer function of the periodic Q1 system (K + lambda^2 M) phi = M n
    (FEM.cpp:73-110 element matrices, square cells of side h): the DFT of the
    9-point stencils, phi_hat = T n_hat."""
    tx = 2 * np.pi * np.fft.fftfreq(n0)
    ty = 2 * np.pi * np.fft.fftfreq(n1)
    cx = np.cos(tx)[None, :]
    cy = np.cos(ty)[:, None]
    K = 8.0 / 3.0 - (2.0 / 3.0) * (cx + cy) - (4.0 / 3.0) * cx * cy
    M = h * h * (2 + cx) * (2 + cy) / 9.0
    return M / (K + lam * lam * M)


def fem_solve_fft(n, h, lam):
    n1, n0 = n.shape
    return np.real(np.fft.ifft2(np.fft.fft2(n) * fem_transfer(n0, n1, h, lam)))

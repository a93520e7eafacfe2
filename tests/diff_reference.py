"""numpy restatement of the differentiable periodic-box B_000 diag chain and
its adjoint, for small meshes (the test-side reference for the engine's
trvg_diff_* entry points).

Forward (compute_bispec_in_gpp_box, threept.cpp:1473-2188, for degrees
(0, 0, 0) and the diag form):
  n(x)   = sum_p w_p W(x - x_p) / V_cell      (field.cpp:569-1112 rules)
  delta  = n - sum(w) / V                      (field.cpp:1235-1243)
  D(k)   = V_cell FFT[delta]
  G(x)   = IFFT[D / (W V)],  F_b(x) = IFFT[shell_b D / W]   (unnormalised
           backward transforms, FFTW sign convention)
  B_b    = norm V_cell / n_b^2 sum_x Re F_b^2 Re G.
The forward is pinned against the oracle's bispectrum of the same catalogue;
the adjoint against central finite differences of the forward.
"""

import numpy as np

ORDERS = {'ngp': 1, 'cic': 2, 'tsc': 3, 'pcs': 4}


def _stencil(pos, n, L, order):
    """Per-axis cell indices and weights (field.cpp:569-1112), and dW/ds."""
    loc = (n * pos) / L
    idx = loc.astype(np.int64)
    s = loc - idx
    if order == 2:
        ix = np.stack([idx, np.where(idx == n - 1, 0, idx + 1)])
        w = np.stack([1. - s, s])
        dw = np.stack([-np.ones_like(s), np.ones_like(s)])
        return ix, w, dw
    raise NotImplementedError("this restatement covers CIC only")


def assign(pos, w, n, L, order=2):
    """Real mesh n(x) (n^3, row-major), with index wrap as the reference."""
    mesh = np.zeros(n ** 3)
    st = [_stencil(pos[:, a], n, L, order) for a in range(3)]
    for i in range(order):
        for j in range(order):
            for k in range(order):
                gid = (st[0][0][i] * n + st[1][0][j]) * n + st[2][0][k]
                val = w * st[0][1][i] * st[1][1][j] * st[2][1][k]
                ok = (gid >= 0) & (gid < n ** 3)
                np.add.at(mesh, gid[ok], val[ok])
    return mesh.reshape(n, n, n) / (L / n) ** 3


def assign_adjoint(gmesh, pos, w, n, L, order=2):
    g = gmesh.reshape(-1)
    st = [_stencil(pos[:, a], n, L, order) for a in range(3)]
    val = np.zeros(len(pos))
    d = np.zeros((len(pos), 3))
    for i in range(order):
        for j in range(order):
            for k in range(order):
                gid = (st[0][0][i] * n + st[1][0][j]) * n + st[2][0][k]
                v = g[gid]
                wi, wj, wk = st[0][1][i], st[1][1][j], st[2][1][k]
                di, dj, dk = st[0][2][i], st[1][2][j], st[2][2][k]
                val += v * wi * wj * wk
                d[:, 0] += v * di * wj * wk
                d[:, 1] += v * wi * dj * wk
                d[:, 2] += v * wi * wj * dk
    vc = (L / n) ** 3
    return (w[:, None] / vc) * d * (n / L), val / vc


def _kgrid(n, L):
    dk = 2 * np.pi / L
    c = np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n) * dk
    kx, ky, kz = np.meshgrid(c, c, c, indexing='ij')
    return np.sqrt((kx * kx + ky * ky) + kz * kz)


def _window(n, order):
    ii = np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n)
    u = np.pi * ii / n
    s = np.where(ii != 0, np.sin(u) / np.where(u == 0, 1., u), 1.)
    w1 = s[:, None, None] * s[None, :, None] * s[None, None, :]
    return w1 ** order


def _bwd(a):
    return np.fft.ifftn(a) * a.size


def forward(delta, L, edges, norm, order=2):
    n = delta.shape[0]
    vc, V = (L / n) ** 3, L ** 3
    D = vc * np.fft.fftn(delta)
    W = _window(n, order)
    km = _kgrid(n, L)
    G = _bwd(D / (W * V)).real
    out, nmodes = [], []
    for b in range(len(edges) - 1):
        sh = (edges[b] <= km) & (km < edges[b + 1])
        nb = int(sh.sum())
        F = _bwd(np.where(sh, D / W, 0.)).real
        out.append(norm * vc * np.sum(F * F * G) / float(nb) ** 2)
        nmodes.append(nb)
    return np.array(out), np.array(nmodes)


def adjoint(delta, L, edges, norm, gB, order=2):
    """dL/ddelta for L = sum_b gB_b B_b."""
    n = delta.shape[0]
    vc, V = (L / n) ** 3, L ** 3
    D = vc * np.fft.fftn(delta)
    W = _window(n, order)
    km = _kgrid(n, L)
    G = _bwd(D / (W * V)).real
    acc = np.zeros(delta.shape, complex)
    gG = np.zeros(delta.shape)
    for b in range(len(edges) - 1):
        sh = (edges[b] <= km) & (km < edges[b + 1])
        nb = float(sh.sum())
        a = gB[b] * norm * vc / nb ** 2
        F = _bwd(np.where(sh, D / W, 0.)).real
        gG += a * F * F
        acc += np.where(sh, np.fft.fftn(2 * a * F * G) / W, 0.)
    acc += np.fft.fftn(gG) / (W * V)
    return vc * _bwd(acc).real

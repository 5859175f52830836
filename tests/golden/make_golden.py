"""Writes the golden fixtures under tests/golden/.

The reference ships no golden vectors (its tests print results for a human
to read), so these fixtures hold what the mathematics fixes: FP64 inputs
and 50-digit mpmath direct sums of the reference's kernels
(src/direct-laplace.c:200-336 with the 1/4pi factor on every term,
src/direct.c:35-294 with G = h0(kr)/(4pi) = e^{ikr}/(4pi i k r) and the
dipole term d n.grad_x G of src/local.c:155-166).
Run: python tests/golden/make_golden.py
"""

import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def laplace():
    rng = np.random.default_rng(2024)
    ns, nt = 12, 6
    xs = rng.uniform(-1, 1, (ns, 3))
    q = rng.normal(size=ns)
    nrm = rng.normal(size=(ns, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=ns)
    xt = rng.uniform(-1, 1, (nt, 3))
    f = []
    for t in xt:
        phi = mp.mpf(0)
        g = [mp.mpf(0)] * 3
        for s in range(ns):
            r = [mp.mpf(t[a]) - mp.mpf(xs[s, a]) for a in range(3)]
            r2 = sum(v * v for v in r)
            rr = mp.sqrt(r2)
            nr = sum(mp.mpf(nrm[s, a]) * r[a] for a in range(3))
            phi += mp.mpf(q[s]) / rr + mp.mpf(d[s]) * nr / rr ** 3
            for a in range(3):
                g[a] += -mp.mpf(q[s]) * r[a] / rr ** 3 + mp.mpf(d[s]) * (
                    mp.mpf(nrm[s, a]) / rr ** 3 - 3 * nr * r[a] / rr ** 5)
        c = 1 / (4 * mp.pi)
        f.append([float(phi * c)] + [float(v * c) for v in g])
    return dict(xs=xs.tolist(), q=q.tolist(), nrm=nrm.tolist(), d=d.tolist(),
                xt=xt.tolist(), f=f)


def helmholtz():
    rng = np.random.default_rng(2025)
    ns, nt, k = 10, 5, 2.3
    xs = rng.uniform(-1, 1, (ns, 3))
    q = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    nrm = rng.normal(size=(ns, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    xt = rng.uniform(-1, 1, (nt, 3))
    kk = mp.mpf(k)
    out = []
    for t in xt:
        pot = mp.mpc(0)
        g = [mp.mpc(0)] * 3
        for s in range(ns):
            rv = [mp.mpf(t[a]) - mp.mpf(xs[s, a]) for a in range(3)]
            r = mp.sqrt(sum(v * v for v in rv))
            G = mp.exp(1j * kk * r) / (4 * mp.pi * 1j * kk * r)
            dG = G * (1j * kk - 1 / r)
            ddG = G * ((1j * kk - 1 / r) ** 2 + 1 / r ** 2)
            qq = mp.mpc(q[s].real, q[s].imag)
            dd = mp.mpc(d[s].real, d[s].imag)
            nr = sum(mp.mpf(nrm[s, a]) * rv[a] for a in range(3)) / r
            pot += qq * G + dd * dG * nr
            for a in range(3):
                ra = rv[a] / r
                g[a] += qq * dG * ra + dd * (ddG * nr * ra + dG * (mp.mpf(nrm[s, a]) - nr * ra) / r)
        out.append([pot] + g)
    return dict(k=k, xs=xs.tolist(), q_re=q.real.tolist(), q_im=q.imag.tolist(),
                nrm=nrm.tolist(), d_re=d.real.tolist(), d_im=d.imag.tolist(), xt=xt.tolist(),
                f_re=[[float(mp.re(v)) for v in row] for row in out],
                f_im=[[float(mp.im(v)) for v in row] for row in out])


if __name__ == "__main__":
    with open(os.path.join(HERE, "laplace_direct.json"), "w") as fh:
        json.dump(laplace(), fh, indent=1)
    with open(os.path.join(HERE, "helmholtz_direct.json"), "w") as fh:
        json.dump(helmholtz(), fh, indent=1)
    print("wrote laplace_direct.json, helmholtz_direct.json")

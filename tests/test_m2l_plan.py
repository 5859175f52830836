"""CPU check of the MFMA M2L plan (include/wbfmm_hip.h, wbfmm_tables.c).

The tables the kernel uses (bin-packed A tiles in lane order, the slot
maps and the per-offset theta/chi/r) are applied here with numpy exactly
as the kernel applies them -- azimuthal rotation Zin, three tile products
with the t^-k scalings, azimuthal rotation Zout -- and the result is
compared with the oracle's rotate / coaxial S|R / rotate-back M2L
(src/pass-laplace.c:73-101) for every list-4 direction class.  A wrong
slot, sign, row/column swap or theta index fails here without a GPU."""

import math

import numpy as np
import pytest


def emulate(P, A1, A3, A2, CS, M, o):
    N, NU, NB = P.N, P.NU, P.NB
    S = P.S + 1              # slot P.S is the dummy destination
    assert P.S == max(16 * NB, 2 * NU)
    th, ri = P.th[o], P.rix[o]
    assert P.rr[ri] == P.r[o]
    x = np.zeros(S)
    for n in range(N + 1):
        for m in range(n + 1):
            u = n * (n + 1) // 2 + m
            re = M[n * n + max(2 * m - 1, 0)]
            im = M[n * n + 2 * m] if m else 0.0
            c, s = CS[o, m]
            x[P.s1[u] & 0xffff] = c * re - s * im
            x[(P.s1[u] >> 16) & 0xffff] = s * re + c * im   # m = 0: dummy slot

    def tiles(A, v):
        y = np.zeros(S)
        for b in range(NB):
            for j in range(P.nk[b]):
                a = A[P.koff[b] + j].reshape(4, 16).T        # a[r][kk] = lane kk*16 + r
                y[16 * b:16 * b + 16] += a @ v[P.kbase[b] + 4 * j:P.kbase[b] + 4 * j + 4]
        return y

    def move(y, mp):
        z = np.zeros(S)
        for r in range(16 * NB):
            z[mp[r] >> 6] = y[r]
        return z

    y = move(tiles(A1[th], x), P.map1)
    z = move(tiles(A2[ri], y), P.map2)
    a = move(tiles(A3[th], z), P.map3)
    L = np.zeros((N + 1) ** 2)
    for n in range(N + 1):
        for nu in range(n + 1):
            u = n * (n + 1) // 2 + nu
            tr, ti = a[u], a[NU + u]
            c, s = CS[o, nu]
            if nu == 0:
                L[n * n] += tr
            else:
                L[n * n + 2 * nu - 1] += c * tr + s * ti
                L[n * n + 2 * nu] += c * ti - s * tr
    return L


@pytest.mark.parametrize("N", [1, 2, 5, 8, 12, 13, 15, 16, 17, 20, 24])
def test_plan_matches_oracle_m2l(wb, oracle, N):
    w = 0.37
    P, A1, A3, A2, CS = wb.m2l_tables(N, w)
    assert P.N == N and P.C == (N + 1) ** 2
    # every bin holds at most 16 rows and its k steps cover its block(s):
    # shared bins read their own rows, the bins of a block above 16 rows
    # read the block's rows from its first one
    assert sum(P.nk[b] for b in range(P.NB)) == P.KT
    for b in range(P.NB):
        assert 1 <= P.nk[b] and P.kbase[b] % 16 == 0 and P.kbase[b] <= 16 * b
        if P.kbase[b] == 16 * b and (b + 1 == P.NB or P.kbase[b + 1] != 16 * b):
            assert P.nk[b] <= 4
        assert P.kbase[b] + 4 * P.nk[b] <= 16 * P.NB
    # the maps are injective on their real rows
    for mp in (P.map1, P.map2, P.map3):
        dest = [v >> 6 for v in mp[:16 * P.NB] if (v >> 6) != P.S]
        assert len(dest) == len(set(dest)) == P.C and max(dest) < P.S
    assert P.NR == 15
    rng = np.random.default_rng(N)
    offs = [(2, 0, 0), (-3, 1, 2), (0, 0, 2), (0, 0, -3), (3, -3, 3), (2, 3, -2), (-2, -2, 0), (0, 3, 1)]
    for off in offs:
        o = (off[0] + 3) * 49 + (off[1] + 3) * 7 + off[2] + 3
        src = rng.uniform(-0.5, 0.5, (5, 3)) * w
        q = rng.normal(size=5)
        M = np.zeros((N + 1) ** 2)
        for s_, qq in zip(src, q):
            oracle.lap_p2m(N, np.zeros(3), s_, qq, M)
        th, ph, ch, r = oracle.shift_angles(*off)
        H = oracle.H_rotation(N, th)
        Lref = oracle.lap_rotate(oracle.lap_coaxial(2, oracle.lap_rotate(M, N, H, ph, ch), N, N, r * w),
                                 N, H, ch, ph)
        L = emulate(P, A1, A3, A2, CS, M, o)
        assert np.abs(L - Lref).max() < 1e-12 * np.abs(Lref).max(), off


def test_orders_beyond_the_plan_are_refused(wb):
    import ctypes
    P = wb.M2LPlan()
    assert wb.lib().wbfmm_hip_lap_m2l_tables(28, ctypes.byref(P), None, None, None) != 0


def test_theta_table_has_49_angles(wb):
    P, _, _, _, CS = wb.m2l_tables(4, 1.0)
    assert P.NTH == 49
    # offsets on the z axis use chi = 0; r is the offset length
    o = (0 + 3) * 49 + (0 + 3) * 7 + (2 + 3)
    assert P.chi[o] == 0.0 and P.r[o] == 2.0
    assert np.allclose(CS[o], [[1.0, 0.0]] * 5)


def emulate_shift(P, A1, A3, A2, CS, th, M):
    """the MFMA M2M / L2L kernels' arithmetic for one vector: Zin, three
    tile products, Zout (same slots and maps as the M2L)"""
    N, NU, NB = P.N, P.NU, P.NB
    S = P.S + 1
    x = np.zeros(S)
    for n in range(N + 1):
        for m in range(n + 1):
            u = n * (n + 1) // 2 + m
            re = M[n * n + max(2 * m - 1, 0)]
            im = M[n * n + 2 * m] if m else 0.0
            c, s = CS[m]
            x[P.s1[u] & 0xffff] = c * re - s * im
            x[(P.s1[u] >> 16) & 0xffff] = s * re + c * im

    def tiles(A, v):
        y = np.zeros(S)
        for b in range(NB):
            for j in range(P.nk[b]):
                a = A[P.koff[b] + j].reshape(4, 16).T
                y[16 * b:16 * b + 16] += a @ v[P.kbase[b] + 4 * j:P.kbase[b] + 4 * j + 4]
        return y

    def move(y, mp):
        z = np.zeros(S)
        for r in range(16 * NB):
            z[mp[r] >> 6] = y[r]
        return z

    a = move(tiles(A3[th], move(tiles(A2, move(tiles(A1[th], x), P.map1)), P.map2)), P.map3)
    out = np.zeros((N + 1) ** 2)
    for n in range(N + 1):
        for nu in range(n + 1):
            u = n * (n + 1) // 2 + nu
            tr, ti = a[u], a[NU + u]
            c, s = CS[nu]
            if nu == 0:
                out[n * n] = tr
            else:
                out[n * n + 2 * nu - 1] = c * tr + s * ti
                out[n * n + 2 * nu] = c * ti - s * tr
    return out


def shift_tables(wb, N, t):
    import ctypes
    P = wb.M2LPlan()
    assert wb.lib().wbfmm_hip_lap_m2l_tables(N, ctypes.byref(P), None, None, None) == 0
    A1 = np.zeros((2, P.KT, 64))
    A3 = np.zeros((2, P.KT, 64))
    CSu = np.zeros((8, N + 1, 2))
    CSd = np.zeros((8, N + 1, 2))
    L = wb.lib()
    L.wbfmm_hip_lap_shift_tables.argtypes = [ctypes.c_void_p] * 5
    L.wbfmm_hip_lap_shift_coax.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_void_p]
    assert L.wbfmm_hip_lap_shift_tables(ctypes.byref(P), wb.ptr(A1), wb.ptr(A3), wb.ptr(CSu),
                                        wb.ptr(CSd)) == 0
    SS = np.zeros((P.KT, 64))
    RR = np.zeros((P.KT, 64))
    assert L.wbfmm_hip_lap_shift_coax(ctypes.byref(P), 0, t, wb.ptr(SS)) == 0
    assert L.wbfmm_hip_lap_shift_coax(ctypes.byref(P), 1, t, wb.ptr(RR)) == 0
    return P, A1, A3, CSu, CSd, SS, RR


@pytest.mark.parametrize("N", [3, 8, 12, 15])
def test_shift_tables_m2m_l2l(wb, oracle, N):
    # M2M: child multipole -> parent equals the multipole of the same
    # sources about the parent centre (the Laplace M2M is exact degree by
    # degree); L2L: parent local expansion -> child gives the same field in
    # the child box (exact for the truncated expansion)
    wc = 0.25
    t = 0.5 * math.sqrt(3.0) * wc
    P, A1, A3, CSu, CSd, SS, RR = shift_tables(wb, N, t)
    rng = np.random.default_rng(N + 40)
    for c in range(8):
        off = (np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], float) - 0.5) * wc
        src = off + rng.uniform(-0.5, 0.5, (4, 3)) * wc
        q = rng.normal(size=4)
        M = np.zeros((N + 1) ** 2)
        for s_, qq in zip(src, q):
            oracle.lap_p2m(N, off, s_, qq, M)
        Mp = emulate_shift(P, A1, A3, SS, CSu[c], (c >> 2) & 1, M)
        Mref = np.zeros((N + 1) ** 2)
        for s_, qq in zip(src, q):
            oracle.lap_p2m(N, np.zeros(3), s_, qq, Mref)
        assert np.abs(Mp - Mref).max() < 1e-12 * np.abs(Mref).max(), c
        # a local expansion about the parent centre from far sources
        far = np.array([2.0, 1.0, -1.5]) + rng.uniform(-0.1, 0.1, (3, 3))
        qf = rng.normal(size=3)
        Mf = np.zeros((N + 1) ** 2)
        xc = np.array([2.0, 1.0, -1.5])
        for s_, qq in zip(far, qf):
            oracle.lap_p2m(N, xc, s_, qq, Mf)
        # M2L from xc to the origin: rotate, S|R, rotate back
        rr = np.linalg.norm(xc)
        the = math.acos(-xc[2] / rr)
        chi = math.atan2(-xc[1], -xc[0])
        H = oracle.H_rotation(N, the)
        Lp = oracle.lap_rotate(oracle.lap_coaxial(2, oracle.lap_rotate(Mf, N, H, -0.5 * math.pi, chi),
                                                  N, N, rr), N, H, chi, -0.5 * math.pi)
        Lc = emulate_shift(P, A1, A3, RR, CSd[c], 1 - ((c >> 2) & 1), Lp)
        xp = off + rng.uniform(-0.4, 0.4, 3) * wc
        fl = oracle.lap_l_eval(N, off, Lc, xp, 1)[0]
        fp = oracle.lap_l_eval(N, np.zeros(3), Lp, xp, 1)[0]
        assert abs(fl - fp) < 1e-12 * abs(fp), (c, fl, fp)


def degree_scale(N, w, power):
    """w^(n + power) for every packed coefficient of degree n <= N"""
    return np.array([w ** (n + power) for n in range(N + 1) for _ in range(2 * n + 1)])


@pytest.mark.parametrize("N", [3, 8, 12, 15])
def test_width_scaled_tables(wb, oracle, N):
    # The FP32 passes keep M~_n = M_n / w^n and L~_n = L_n w^(n+1) with
    # their own tables: M2L at unit width and level-independent M2M / L2L.
    # Applied to scaled vectors they must give the scaled results of the
    # FP64 tables (a wrong power of 2, sign or exponent fails).
    import ctypes
    wc = 0.25
    wp = 2 * wc
    t = 0.5 * math.sqrt(3.0) * wc
    P, A1, A3, CSu, CSd, SS, RR = shift_tables(wb, N, t)
    L = wb.lib()
    L.wbfmm_hip_lap_shift_coax_unit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    SSu = np.zeros((P.KT, 64))
    RRu = np.zeros((P.KT, 64))
    assert L.wbfmm_hip_lap_shift_coax_unit(ctypes.byref(P), 0, wb.ptr(SSu)) == 0
    assert L.wbfmm_hip_lap_shift_coax_unit(ctypes.byref(P), 1, wb.ptr(RRu)) == 0
    rng = np.random.default_rng(N + 90)
    for c in (0, 3, 5, 7):
        M = np.zeros((N + 1) ** 2)
        off = (np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], float) - 0.5) * wc
        for s_ in off + rng.uniform(-0.5, 0.5, (3, 3)) * wc:
            oracle.lap_p2m(N, off, s_, rng.normal(), M)
        Mp = emulate_shift(P, A1, A3, SS, CSu[c], (c >> 2) & 1, M)
        Mpu = emulate_shift(P, A1, A3, SSu, CSu[c], (c >> 2) & 1, M / degree_scale(N, wc, 0))
        assert np.abs(Mpu * degree_scale(N, wp, 0) - Mp).max() < 1e-12 * np.abs(Mp).max(), c
        Lp = rng.normal(size=(N + 1) ** 2) / degree_scale(N, wp, 1)
        Lc = emulate_shift(P, A1, A3, RR, CSd[c], 1 - ((c >> 2) & 1), Lp)
        Lcu = emulate_shift(P, A1, A3, RRu, CSd[c], 1 - ((c >> 2) & 1),
                            Lp * degree_scale(N, wp, 1))
        assert np.abs(Lcu / degree_scale(N, wc, 1) - Lc).max() < 1e-12 * np.abs(Lc).max(), c
    # M2L: unit-width S|R on M~ gives L~
    w = 0.37
    Pm, B1, B3, B2, BCS = wb.m2l_tables(N, w)
    _, _, _, B2u, _ = wb.m2l_tables(N, 1.0)
    for off in [(2, 0, 0), (-3, 1, 2), (3, -3, 3)]:
        o = (off[0] + 3) * 49 + (off[1] + 3) * 7 + off[2] + 3
        M = rng.normal(size=(N + 1) ** 2) * degree_scale(N, w, 0)
        Lw = emulate(Pm, B1, B3, B2, BCS, M, o)
        Lu = emulate(Pm, B1, B3, B2u, BCS, M / degree_scale(N, w, 0), o)
        assert np.abs(Lu / degree_scale(N, w, 1) - Lw).max() < 1e-12 * np.abs(Lw).max(), off


def emulate_hel_core(P, T1, Ar, Ai, chi, M, N):
    """one Helmholtz translation as the kernels apply the tiles: phase
    e^{i m chi} in, P/Q basis, the rotation tiles T1, the coaxial tiles
    (real part Ar, imaginary part Ai or None), T1 again, back to x_{+-nu}
    and the phase e^{-i nu chi} out.  M and the result: complex, index
    n(n+1)+m."""
    NU, NB = P.NU, P.NB
    S = P.S + 1
    x = np.zeros(S, complex)
    for n in range(N + 1):
        for m in range(n + 1):
            u = n * (n + 1) // 2 + m
            e = np.exp(1j * m * chi)
            xp, xm = e * M[n * (n + 1) + m], np.conj(e) * M[n * (n + 1) - m]
            if m == 0:
                x[P.s1[u] & 0xffff] = 2 * xp
            else:
                x[P.s1[u] & 0xffff] = xp + xm
                x[(P.s1[u] >> 16) & 0xffff] = xp - xm

    def tiles(T, v):
        y = np.zeros(S, complex)
        for b in range(NB):
            for j in range(P.nk[b]):
                a = T[P.koff[b] + j].reshape(4, 16).T
                y[16 * b:16 * b + 16] += a @ v[P.kbase[b] + 4 * j:P.kbase[b] + 4 * j + 4]
        return y

    def move(y, mp):
        z = np.zeros(S, complex)
        for r in range(16 * NB):
            z[mp[r] >> 6] = y[r]
        return z

    y = move(tiles(T1, x), P.map1)
    z = tiles(Ar, y)
    if Ai is not None:
        z = z + 1j * tiles(Ai, y)
    z = move(z, P.map2)
    a = move(tiles(T1, z), P.map3)
    L = np.zeros((N + 1) ** 2, complex)
    for n in range(N + 1):
        for nu in range(n + 1):
            u = n * (n + 1) // 2 + nu
            pp, qq = a[u], a[NU + u]
            if nu == 0:
                L[n * n + n] = 0.5 * pp
            else:
                L[n * n + n + nu] = np.exp(-1j * nu * chi) * 0.5 * (pp + qq)
                L[n * n + n - nu] = np.exp(1j * nu * chi) * 0.5 * (pp - qq)
    return L


def emulate_hel(P, A, A2, CS, M, o, N):
    """the Helmholtz M2L of list-4 slot o"""
    assert np.allclose(CS[o, :, 0], np.cos(np.arange(N + 1) * P.chi[o]))
    return emulate_hel_core(P, A[P.th[o]], A2[P.rix[o], 0], A2[P.rix[o], 1], P.chi[o], M, N)


@pytest.mark.parametrize("N,L", [(2, 2), (5, 7), (9, 9), (13, 13), (16, 18), (20, 20), (27, 27)])
def test_hel_plan_matches_oracle_m2l(wb, oracle, N, L):
    # the Helmholtz M2L tiles over the Laplace plan of the same order
    # against the oracle's rotate / coaxial S|R / rotate back
    # (src/pass.c:56-105) for several list-4 directions; the S|R tables
    # are of order L >= N as when the source and receiver orders differ
    k, w = 2.3, 0.41
    P0, _, _, _, CS = wb.m2l_tables(N, w)
    ne = (L + 1) * (L + 2) * (L + 3) // 6
    E = np.stack([oracle.hel_coaxial_table(1, L, k * w * P0.rr[i])[:2 * ne] for i in range(P0.NR)])
    P, A, A2 = wb.hel_m2l_tables(N, L, E)
    assert P.NTH == 49 and A2.shape == (15, 2, P.KT, 64)
    rng = np.random.default_rng(100 + N)
    offs = [(2, 0, 0), (-3, 1, 2), (0, 0, 2), (0, 0, -3), (3, -3, 3), (2, 3, -2), (-2, -2, 0), (0, 3, 1)]
    for off in offs:
        o = (off[0] + 3) * 49 + (off[1] + 3) * 7 + off[2] + 3
        M = np.zeros(2 * (N + 1) ** 2)
        for _ in range(3):
            s_ = rng.uniform(-0.5, 0.5, 3) * w
            oracle.hel_p2m(k, N, np.zeros(3), s_, complex(*rng.normal(size=2)), M)
        th, ph, ch, r = oracle.shift_angles(*off)
        H = oracle.H_rotation(N, th)
        Ef = oracle.hel_coaxial_table(1, L, k * w * r)
        Lref = oracle.hel_rotate(oracle.hel_coaxial(oracle.hel_rotate(M, N, H, ph, ch), N, N, Ef, L, 1),
                                 N, H, ch, ph)
        Lref = Lref[0::2] + 1j * Lref[1::2]
        Lc = emulate_hel(P, A, A2, CS, M[0::2] + 1j * M[1::2], o, N)
        assert np.abs(Lc - Lref).max() < 1e-12 * np.abs(Lref).max(), off


def child_angles(c, sgn):
    ex, ey, ez = (sgn * (0.5 - ((c >> i) & 1)) for i in range(3))
    return math.acos(ez / math.sqrt(ex * ex + ey * ey + ez * ez)), -0.5 * math.pi, math.atan2(ey, ex)


@pytest.mark.parametrize("N", [2, 7, 12, 16, 23, 26])
def test_hel_shift_tiles_match_oracle(wb, oracle, N):
    # Helmholtz M2M (child -> parent) and L2L (parent -> child): the rotation
    # tiles at the two child polar angles, the real R|R coaxial tiles at
    # sqrt(3)/2 of the child width and the child-direction phases of the
    # Laplace shift tables, against the oracle's rotate / coaxial R|R /
    # rotate back (src/pass.c:550-632)
    import ctypes
    k, wc = 3.7, 0.23
    P, _, _, CSu, CSd, _, _ = shift_tables(wb, N, 1.0)
    ths = np.array([math.acos(1 / math.sqrt(3)), math.acos(-1 / math.sqrt(3))])
    A = np.zeros((2, P.KT, 64))
    assert wb.lib().wbfmm_hip_hel_rot_tiles(ctypes.byref(P), 2, wb.ptr(ths), wb.ptr(A)) == 0
    ne = (N + 1) * (N + 2) * (N + 3) // 6
    E = oracle.hel_coaxial_table(0, N, k * wc * 0.5 * math.sqrt(3.0))
    _, _, A2 = wb.hel_m2l_tables(N, N, E[None, :2 * ne])
    rng = np.random.default_rng(200 + N)
    for mode, CS, sgn in ((0, CSu, 1), (1, CSd, -1)):
        M = np.zeros(2 * (N + 1) ** 2)
        for _ in range(3):
            s_ = rng.uniform(-0.5, 0.5, 3) * wc
            oracle.hel_p2m(k, N, np.zeros(3), s_, complex(*rng.normal(size=2)), M)
        for c in range(8):
            th, ph, ch = child_angles(c, sgn)
            assert np.allclose(CS[c, :, 0], np.cos(np.arange(N + 1) * ch))
            assert np.allclose(CS[c, :, 1], np.sin(np.arange(N + 1) * ch))
            ti = (c >> 2) & 1 if mode == 0 else 1 - ((c >> 2) & 1)
            assert abs(ths[ti] - th) < 1e-14
            H = oracle.H_rotation(N, th)
            Lref = oracle.hel_rotate(oracle.hel_coaxial(oracle.hel_rotate(M, N, H, ph, ch), N, N, E,
                                                        N, 0), N, H, ch, ph)
            Lref = Lref[0::2] + 1j * Lref[1::2]
            Lc = emulate_hel_core(P, A[ti], A2[0, 0], None, ch, M[0::2] + 1j * M[1::2], N)
            assert np.abs(Lc - Lref).max() < 1e-12 * np.abs(Lref).max(), (mode, c)

"""CPU restatement of maxmin (src/maxmin.F:5-119) for small arrays (test
infrastructure only): per level, the first point in j, then i order that
beats the running extreme (strict comparison, starting from -1e30 / 1e30),
then the first level in k order that beats the others."""


def maxmin(f, izz, jzz, kzz):
    """f indexed f[k][j][i] over k, j, i = 1..kzz, jzz, izz (index 0 unused
    or halo); returns (fmax, (i, j, k)), (fmin, (i, j, k))."""
    tmax, tmin, lmax, lmin = [], [], [], []
    for k in range(1, kzz + 1):
        a, b, la, lb = -1.0e30, 1.0e30, None, None
        for j in range(1, jzz + 1):
            for i in range(1, izz + 1):
                v = float(f[k][j][i])
                if v > a:
                    a, la = v, (i, j, k)
                if v < b:
                    b, lb = v, (i, j, k)
        tmax.append(a), tmin.append(b), lmax.append(la), lmin.append(lb)
    fmax, fmin, imax, imin = -1.0e30, 1.0e30, (1, 1, 1), (1, 1, 1)
    for k in range(kzz):
        if tmax[k] > fmax:
            fmax, imax = tmax[k], lmax[k]
        if tmin[k] < fmin:
            fmin, imin = tmin[k], lmin[k]
    return (fmax, imax), (fmin, imin)

/* orc_guided_impl.h -- restatement of guided.c, instantiated twice by
 * oracle_guided.c (REAL=float: the reference arithmetic; REAL=double: fp64 twin).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */

/* separable (2r+1)^3 mean, edge windows shrink to the in-bounds count, sliding
 * REAL sums along x, then y, then z with per-axis 1/count factors (guided.c:22-98) */
void SFX(orc_box_mean)(const REAL *in, REAL *out, int nz, int ny, int nx, int r) {
    size_t plane = (size_t)ny * nx, n = plane * nz;
    REAL *t1 = (REAL *)malloc(sizeof(REAL) * n), *t2 = (REAL *)malloc(sizeof(REAL) * n);
    int lm = nx > ny ? nx : ny; if (nz > lm) lm = nz;
    REAL *rc = (REAL *)malloc(sizeof(REAL) * (size_t)lm * 3);
    REAL *rcx = rc, *rcy = rc + lm, *rcz = rc + 2 * lm;
    int dims[3] = {nx, ny, nz};
    REAL *rcs[3] = {rcx, rcy, rcz};
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < dims[a]; i++) {
            int lo = i - r < 0 ? 0 : i - r, hi = i + r > dims[a] - 1 ? dims[a] - 1 : i + r;
            rcs[a][i] = (REAL)1 / (REAL)(hi - lo + 1);
        }
    /* one sliding pass along an axis with element stride st, length len */
    const REAL *src = in; REAL *dst = t1;
    for (int a = 0; a < 3; a++) {
        int len = dims[a];
        size_t st = a == 0 ? 1 : (a == 1 ? (size_t)nx : plane);
        for (size_t line = 0; line < n / (size_t)len; line++) {
            size_t base;
            if (a == 0) base = line * nx;
            else if (a == 1) base = (line / nx) * plane + (line % nx);
            else base = line;
            REAL acc = 0;
            for (int i = 0; i <= r && i < len; i++) acc += src[base + i * st];
            for (int i = 0; i < len; i++) {
                dst[base + i * st] = acc * rcs[a][i];
                if (i + r + 1 < len) acc += src[base + (size_t)(i + r + 1) * st];
                if (i - r >= 0) acc -= src[base + (size_t)(i - r) * st];
            }
        }
        src = dst; dst = (a == 0) ? t2 : out;
    }
    free(t1); free(t2); free(rc);
}

/* self-guided filter, guide == input (guided.c:113-137) */
int SFX(orc_guided_exact)(const REAL *in, REAL *out, int nz, int ny, int nx, int r, double eps) {
    if (r < 1) r = 2;
    size_t n = (size_t)nz * ny * nx;
    REAL *m = (REAL *)malloc(sizeof(REAL) * n), *q = (REAL *)malloc(sizeof(REAL) * n);
    REAL *a = (REAL *)malloc(sizeof(REAL) * n), *b = (REAL *)malloc(sizeof(REAL) * n);
    if (!m || !q || !a || !b) { free(m); free(q); free(a); free(b); return 1; }
    SFX(orc_box_mean)(in, m, nz, ny, nx, r);
    for (size_t i = 0; i < n; i++) a[i] = in[i] * in[i];
    SFX(orc_box_mean)(a, q, nz, ny, nx, r);
    for (size_t i = 0; i < n; i++) {
        REAL v = q[i] - m[i] * m[i];
        if (v < 0) v = 0;
        REAL ai = v / (v + (REAL)eps);
        a[i] = ai;
        b[i] = m[i] * ((REAL)1 - ai);
    }
    SFX(orc_box_mean)(a, m, nz, ny, nx, r);
    SFX(orc_box_mean)(b, q, nz, ny, nx, r);
    for (size_t i = 0; i < n; i++) out[i] = m[i] * in[i] + q[i];
    free(m); free(q); free(a); free(b);
    return 0;
}

/* He & Sun fast variant: coefficients on the s-decimated grid, trilinear
 * upsample, apply at full resolution (guided.c:152-231) */
int SFX(orc_guided_fast)(const REAL *in, REAL *out, int nz, int ny, int nx, int r, double eps, int s) {
    if (r < 1) r = 2;
    if (s <= 1) return SFX(orc_guided_exact)(in, out, nz, ny, nx, r, eps);
    int lz = (nz + s - 1) / s, ly = (ny + s - 1) / s, lx = (nx + s - 1) / s;
    if (lz < 2 || ly < 2 || lx < 2) return SFX(orc_guided_exact)(in, out, nz, ny, nx, r, eps);
    size_t nl = (size_t)lz * ly * lx;
    REAL *il = (REAL *)malloc(sizeof(REAL) * nl), *ma = (REAL *)malloc(sizeof(REAL) * nl);
    REAL *mb = (REAL *)malloc(sizeof(REAL) * nl);
    for (int z = 0; z < lz; z++)
        for (int y = 0; y < ly; y++)
            for (int x = 0; x < lx; x++)
                il[((size_t)z * ly + y) * lx + x] = in[((size_t)(z * s) * ny + (size_t)(y * s)) * nx + (size_t)x * s];
    int rl = r / s; if (rl < 1) rl = 1;
    /* coefficient field: the exact filter's a/b stage on the low-res grid */
    REAL *m = (REAL *)malloc(sizeof(REAL) * nl), *q = (REAL *)malloc(sizeof(REAL) * nl);
    REAL *a = (REAL *)malloc(sizeof(REAL) * nl), *b = (REAL *)malloc(sizeof(REAL) * nl);
    SFX(orc_box_mean)(il, m, lz, ly, lx, rl);
    for (size_t i = 0; i < nl; i++) a[i] = il[i] * il[i];
    SFX(orc_box_mean)(a, q, lz, ly, lx, rl);
    for (size_t i = 0; i < nl; i++) {
        REAL v = q[i] - m[i] * m[i];
        if (v < 0) v = 0;
        REAL ai = v / (v + (REAL)eps);
        a[i] = ai;
        b[i] = m[i] * ((REAL)1 - ai);
    }
    SFX(orc_box_mean)(a, ma, lz, ly, lx, rl);
    SFX(orc_box_mean)(b, mb, lz, ly, lx, rl);
    free(m); free(q); free(a); free(b);
    /* trilinear: u = coord/s, i0 = min((int)u, l-2), frac = min(u-i0, 1);
     * z/y blend of the four rows first, then the x lerp (guided.c:193-228) */
    REAL *arow = (REAL *)malloc(sizeof(REAL) * (size_t)lx), *brow = (REAL *)malloc(sizeof(REAL) * (size_t)lx);
    for (int z = 0; z < nz; z++) {
        float uz = (float)z / s; int z0 = (int)uz; if (z0 > lz - 2) z0 = lz - 2;
        float fzr = uz - z0; if (fzr > 1.0f) fzr = 1.0f;
        REAL zf = (REAL)fzr;
        for (int y = 0; y < ny; y++) {
            float uy = (float)y / s; int y0 = (int)uy; if (y0 > ly - 2) y0 = ly - 2;
            float fyr = uy - y0; if (fyr > 1.0f) fyr = 1.0f;
            REAL yf = (REAL)fyr;
            size_t q00 = ((size_t)z0 * ly + y0) * lx, q01 = q00 + lx;
            size_t q10 = q00 + (size_t)ly * lx, q11 = q10 + lx;
            REAL w00 = (1 - zf) * (1 - yf), w01 = (1 - zf) * yf, w10 = zf * (1 - yf), w11 = zf * yf;
            for (int k = 0; k < lx; k++) {
                arow[k] = w00 * ma[q00 + k] + w01 * ma[q01 + k] + w10 * ma[q10 + k] + w11 * ma[q11 + k];
                brow[k] = w00 * mb[q00 + k] + w01 * mb[q01 + k] + w10 * mb[q10 + k] + w11 * mb[q11 + k];
            }
            const REAL *irow = in + ((size_t)z * ny + y) * nx;
            REAL *orow = out + ((size_t)z * ny + y) * nx;
            for (int x = 0; x < nx; x++) {
                float ux = (float)x / s; int x0 = (int)ux; if (x0 > lx - 2) x0 = lx - 2;
                float fxr = ux - x0; if (fxr > 1.0f) fxr = 1.0f;
                REAL fx = (REAL)fxr, gx = 1 - fx;
                REAL av = gx * arow[x0] + fx * arow[x0 + 1];
                REAL bv = gx * brow[x0] + fx * brow[x0 + 1];
                orow[x] = av * irow[x] + bv;
            }
        }
    }
    free(arow); free(brow); free(il); free(ma); free(mb);
    return 0;
}

/* orc_deconv_impl.h -- restatement of deconv_radial (paganin.c:152-293),
 * instantiated twice by oracle_deconv.c:
 *   REAL=float  -> orc_deconv_core_f32 (the reference arithmetic)
 *   REAL=double -> orc_deconv_core_f64 (fp64 twin that keeps the LUT semantics:
 *                  the 4097-entry table over f^2 with linear interpolation, but
 *                  with entries, FFTs and products in double)
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Steps: reflect-padding to fy_next_fft_size per axis; forward X as a two-rows-per-
 * complex-FFT real transform untangled to the half spectrum kx in [0, px/2];
 * forward Y and Z on the half spectrum; multiply by the interpolated LUT weight at
 * u = fz^2*s + fy^2 + fx^2; inverse Z (all rows), inverse Y (z < nz), inverse X as
 * complex-to-real pairs (y < ny), scale 1/(pz*py*px), crop.
 */

static int SFX(orc_deconv_core)(const float *in, REAL *out, int nz, int ny, int nx,
                               const double *lutd, double us, const double *fz2,
                               const double *fy2, const double *fx2) {
    int pz = orc_next_fft_size(nz), py = orc_next_fft_size(ny), px = orc_next_fft_size(nx);
    int hx = px / 2 + 1;
    size_t nh = (size_t)pz * py * hx;
    REAL *sr = (REAL *)malloc(sizeof(REAL) * nh), *si = (REAL *)malloc(sizeof(REAL) * nh);
    int mx = px > py ? px : py; if (pz > mx) mx = pz;
    REAL *lr = (REAL *)malloc(sizeof(REAL) * (size_t)mx), *li = (REAL *)malloc(sizeof(REAL) * (size_t)mx);
    REAL lut[ORC_LUT_N + 1];
    if (!sr || !si || !lr || !li) { free(sr); free(si); free(lr); free(li); return 1; }
    for (int i = 0; i <= ORC_LUT_N; i++) lut[i] = (REAL)lutd[i];

    /* forward X: rows (y, y+1) packed as re/im, reflect gather (paganin.c:166-194) */
    for (int z = 0; z < pz; z++) {
        int rz = orc_reflect(z, nz);
        for (int y = 0; y < py; y += 2) {
            int pair = (y + 1 < py);
            const float *ra = in + ((size_t)rz * ny + orc_reflect(y, ny)) * nx;
            const float *rb = pair ? in + ((size_t)rz * ny + orc_reflect(y + 1, ny)) * nx : NULL;
            for (int x = 0; x < px; x++) {
                int sx = orc_reflect(x, nx);
                lr[x] = (REAL)ra[sx];
                li[x] = pair ? (REAL)rb[sx] : (REAL)0;
            }
            SFX(orc_fft1d)(lr, li, px, -1);
            REAL *ar = sr + ((size_t)z * py + y) * hx, *ai = si + ((size_t)z * py + y) * hx;
            for (int k = 0; k < hx; k++) {
                int nk = k ? px - k : 0;
                ar[k] = (REAL)0.5 * (lr[k] + lr[nk]);
                ai[k] = (REAL)0.5 * (li[k] - li[nk]);
                if (pair) {
                    ar[hx + k] = (REAL)0.5 * (li[k] + li[nk]);   /* row y+1 real */
                    ai[hx + k] = (REAL)0.5 * (lr[nk] - lr[k]);   /* row y+1 imag */
                }
            }
        }
    }
    /* forward Y over every z, then Z over every y (paganin.c:197-210) */
    size_t zs = (size_t)py * hx;
    for (int z = 0; z < pz; z++)
        for (int k = 0; k < hx; k++) {
            size_t b = (size_t)z * zs + k;
            for (int y = 0; y < py; y++) { lr[y] = sr[b + (size_t)y * hx]; li[y] = si[b + (size_t)y * hx]; }
            SFX(orc_fft1d)(lr, li, py, -1);
            for (int y = 0; y < py; y++) { sr[b + (size_t)y * hx] = lr[y]; si[b + (size_t)y * hx] = li[y]; }
        }
    for (size_t c = 0; c < zs; c++) {
        for (int z = 0; z < pz; z++) { lr[z] = sr[c + (size_t)z * zs]; li[z] = si[c + (size_t)z * zs]; }
        SFX(orc_fft1d)(lr, li, pz, -1);
        for (int z = 0; z < pz; z++) { sr[c + (size_t)z * zs] = lr[z]; si[c + (size_t)z * zs] = li[z]; }
    }
    /* LUT weight (paganin.c:231-244) */
    for (int z = 0; z < pz; z++)
        for (int y = 0; y < py; y++) {
            double fzy = fz2[z] + fy2[y];
            size_t row = ((size_t)z * py + y) * hx;
            for (int k = 0; k < hx; k++) {
                double t = (fzy + fx2[k]) * us;
                int ti = (int)t;
                if (ti >= ORC_LUT_N) ti = ORC_LUT_N - 1;
                REAL w = lut[ti] + (lut[ti + 1] - lut[ti]) * (REAL)(t - ti);
                sr[row + k] *= w;
                si[row + k] *= w;
            }
        }
    /* inverse Z (all y), inverse Y (z < nz) (paganin.c:247-260) */
    for (size_t c = 0; c < zs; c++) {
        for (int z = 0; z < pz; z++) { lr[z] = sr[c + (size_t)z * zs]; li[z] = si[c + (size_t)z * zs]; }
        SFX(orc_fft1d)(lr, li, pz, +1);
        for (int z = 0; z < pz; z++) { sr[c + (size_t)z * zs] = lr[z]; si[c + (size_t)z * zs] = li[z]; }
    }
    for (int z = 0; z < nz; z++)
        for (int k = 0; k < hx; k++) {
            size_t b = (size_t)z * zs + k;
            for (int y = 0; y < py; y++) { lr[y] = sr[b + (size_t)y * hx]; li[y] = si[b + (size_t)y * hx]; }
            SFX(orc_fft1d)(lr, li, py, +1);
            for (int y = 0; y < py; y++) { sr[b + (size_t)y * hx] = lr[y]; si[b + (size_t)y * hx] = li[y]; }
        }
    /* inverse X: L = A + iB with Hermitian extension, scale, crop (paganin.c:263-289) */
    REAL scale = (REAL)1 / ((REAL)pz * (REAL)py * (REAL)px);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y += 2) {
            int pair = (y + 1 < ny);
            const REAL *ar = sr + ((size_t)z * py + y) * hx, *ai = si + ((size_t)z * py + y) * hx;
            const REAL *br = ar + hx, *bi = ai + hx;
            for (int k = 0; k < px; k++) {
                int m = k < hx ? k : px - k;
                REAL bre = pair ? br[m] : (REAL)0, bim = pair ? bi[m] : (REAL)0;
                if (k < hx) { lr[k] = ar[m] - bim; li[k] = ai[m] + bre; }
                else        { lr[k] = ar[m] + bim; li[k] = -ai[m] + bre; }
            }
            SFX(orc_fft1d)(lr, li, px, +1);
            REAL *oa = out + ((size_t)z * ny + y) * nx;
            for (int x = 0; x < nx; x++) {
                oa[x] = lr[x] * scale;
                if (pair) oa[nx + x] = li[x] * scale;
            }
        }
    free(sr); free(si); free(lr); free(li);
    return 0;
}

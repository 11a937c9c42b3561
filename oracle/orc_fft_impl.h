/* orc_fft_impl.h -- restatement of fft.c, instantiated twice by oracle_fft.c:
 *   REAL=float  SFX(x)=x##_f32   (the reference's arithmetic)
 *   REAL=double SFX(x)=x##_f64   (the fp64 twin)
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Algorithm (fft.c:47-173): twiddles cos/sin(2*pi*k/len) evaluated in double and
 * stored as REAL, positive angle, sign applied to the sine; bit-reversal then
 * in-place radix-2 decimation-in-time stages len=2..n; n = 3m is split into the
 * three residue sub-sequences x[3j+c], each transformed as a pow2 FFT of length m,
 * then merged with w^k, w^2k and the cube roots of unity.
 */

/* radix-2 in-place transform of a pow2 length (fft.c:78-103) */
static void SFX(pow2_fft)(REAL *re, REAL *im, int n, int sign) {
    if (n <= 1) return;
    /* bit reversal (fft.c:64-75) */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        while (j & bit) { j ^= bit; bit >>= 1; }
        j |= bit;
        if (i < j) {
            REAL t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    REAL *wr = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    REAL *wi = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    const REAL sg = (REAL)sign;
    for (int len = 2; len <= n; len *= 2) {
        int half = len / 2;
        double step = 2.0 * M_PI / (double)len;
        for (int k = 0; k < half; k++) {       /* table entry per stage (fft.c:51-58) */
            wr[k] = (REAL)cos(step * k);
            wi[k] = (REAL)sin(step * k);
        }
        for (int base = 0; base < n; base += len) {
            for (int k = 0; k < half; k++) {
                REAL c = wr[k], s = sg * wi[k];
                int a = base + k, b = a + half;
                REAL xr = re[b] * c - im[b] * s;
                REAL xi = re[b] * s + im[b] * c;
                REAL ur = re[a], ui = im[a];
                re[a] = ur + xr; im[a] = ui + xi;
                re[b] = ur - xr; im[b] = ui - xi;
            }
        }
    }
    free(wr); free(wi);
}

/* n = 3*m split (fft.c:114-166) */
static void SFX(radix3_fft)(REAL *re, REAL *im, int n, int sign) {
    int m = n / 3;
    REAL *s = (REAL *)malloc(sizeof(REAL) * (size_t)6 * m);
    REAL *r0 = s, *i0 = s + m, *r1 = s + 2 * m, *i1 = s + 3 * m, *r2 = s + 4 * m, *i2 = s + 5 * m;
    for (int j = 0; j < m; j++) {
        r0[j] = re[3 * j];     i0[j] = im[3 * j];
        r1[j] = re[3 * j + 1]; i1[j] = im[3 * j + 1];
        r2[j] = re[3 * j + 2]; i2[j] = im[3 * j + 2];
    }
    SFX(pow2_fft)(r0, i0, m, sign);
    SFX(pow2_fft)(r1, i1, m, sign);
    SFX(pow2_fft)(r2, i2, m, sign);
    const REAL sg = (REAL)sign;
    const REAL cr = (REAL)-0.5;
    const REAL ci = sg * (REAL)0.86602540378443864676;   /* e^{sign 2 pi i/3} */
    double step = 2.0 * M_PI / (double)n;
    for (int k = 0; k < m; k++) {
        REAL w1r = (REAL)cos(step * k),     w1i = sg * (REAL)sin(step * k);
        REAL w2r = (REAL)cos(step * 2 * k), w2i = sg * (REAL)sin(step * 2 * k);
        REAL ar = r0[k], ai = i0[k];
        REAL br = r1[k] * w1r - i1[k] * w1i, bi = r1[k] * w1i + i1[k] * w1r;
        REAL dr = r2[k] * w2r - i2[k] * w2i, di = r2[k] * w2i + i2[k] * w2r;
        re[k] = ar + br + dr;
        im[k] = ai + bi + di;
        /* u*b + conj(u)*d */
        REAL p1r = cr * br - ci * bi, p1i = cr * bi + ci * br;
        REAL p2r = cr * dr + ci * di, p2i = cr * di - ci * dr;
        re[k + m] = ar + p1r + p2r;
        im[k + m] = ai + p1i + p2i;
        /* conj(u)*b + u*d */
        REAL q1r = cr * br + ci * bi, q1i = cr * bi - ci * br;
        REAL q2r = cr * dr - ci * di, q2i = cr * di + ci * dr;
        re[k + 2 * m] = ar + q1r + q2r;
        im[k + 2 * m] = ai + q1i + q2i;
    }
    free(s);
}

void SFX(orc_fft1d)(REAL *re, REAL *im, int n, int sign) {
    if (n % 3 == 0 && orc_is_pow2(n / 3)) SFX(radix3_fft)(re, im, n, sign);
    else SFX(pow2_fft)(re, im, n, sign);
}

/* axis order X, Y, Z through a gathered line (fft.c:178-220) */
void SFX(orc_fft3d)(REAL *re, REAL *im, int nz, int ny, int nx, int sign) {
    int mx = nx > ny ? nx : ny; if (nz > mx) mx = nz;
    REAL *lr = (REAL *)malloc(sizeof(REAL) * (size_t)mx);
    REAL *li = (REAL *)malloc(sizeof(REAL) * (size_t)mx);
    size_t plane = (size_t)ny * nx;
    for (size_t row = 0; row < (size_t)nz * ny; row++)
        SFX(orc_fft1d)(re + row * nx, im + row * nx, nx, sign);
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
            size_t b = (size_t)z * plane + x;
            for (int y = 0; y < ny; y++) { lr[y] = re[b + (size_t)y * nx]; li[y] = im[b + (size_t)y * nx]; }
            SFX(orc_fft1d)(lr, li, ny, sign);
            for (int y = 0; y < ny; y++) { re[b + (size_t)y * nx] = lr[y]; im[b + (size_t)y * nx] = li[y]; }
        }
    for (size_t c = 0; c < plane; c++) {
        for (int z = 0; z < nz; z++) { lr[z] = re[c + (size_t)z * plane]; li[z] = im[c + (size_t)z * plane]; }
        SFX(orc_fft1d)(lr, li, nz, sign);
        for (int z = 0; z < nz; z++) { re[c + (size_t)z * plane] = lr[z]; im[c + (size_t)z * plane] = li[z]; }
    }
    free(lr); free(li);
}

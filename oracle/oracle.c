/*
 * CPU oracle for the real-space grid Hamiltonian path -- TEST INFRASTRUCTURE.
 *
 * This is a plain-C, FP64 restatement of the reference's serial algorithms:
 * halo fill, FD stencil, restrict / interpolate, cubic radial splines,
 * spline_to_grid with solid spherical harmonics, the LFC boundary sweep
 * with integrate / add, and the multigrid preconditioner sequence.  It
 * follows the reference's loop order and parenthesisation so that, built
 * with -ffp-contract=off and no FMA-enabling -march, every function is
 * bit-for-bit comparable with the reference's own C code.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this library, and only as the checker.  The product path never calls
 * it.
 *
 * Parity status (how each function is pinned):
 *   orc_halo_fill, orc_fd     -- reference Operator (oracle/_ref) bit-for-bit,
 *                                FD symbol closed form, polynomial exactness.
 *   orc_transform_z           -- reference complex Transformer with phases
 *                                (oracle/_ref) bit-for-bit.
 *   orc_halo_fill_z, orc_fdz  -- reference complex Operator with phases
 *                                (oracle/_ref) bit-for-bit; Bloch plane waves
 *                                are eigenvectors with the FD symbol at k+G.
 *   orc_restrict/interpolate  -- reference Transformer bit-for-bit, integral
 *                                conservation, polynomial reproduction.
 *   orc_spline_*              -- reference Spline bit-for-bit, interpolation
 *                                of the knots, natural end conditions.
 *   orc_spline_to_grid        -- reference spline_to_grid bit-for-bit, closed
 *                                form of solid harmonics.
 *   orc_lfc_integrate/add     -- reference LFC bit-for-bit, adjointness,
 *                                dense brute force on tiny grids.
 *   orc_lfc_integrate_z/add_z -- reference complex LFC with phase_qW
 *                                (oracle/_ref) bit-for-bit; adjointness with
 *                                the phases, which fixes add's conjugate.
 *   orc_precondition          -- composition of the pinned pieces; the final
 *                                axpy uses two roundings (no FMA); the
 *                                reference calls BLAS daxpy there, whose FMA
 *                                use depends on the BLAS build: that one step
 *                                is pinned to 1 ulp, not bitwise.
 *   orc_relax                 -- reference Operator.relax (oracle/_ref)
 *                                bit-for-bit, both methods (golden digests).
 *   orc_add_local_potential   -- the reference's numpy expression
 *                                Htpsit += psit * vt, bit-for-bit.
 *   orc_projector_coefs       -- numpy P @ dH with a non-symmetric dH, to
 *                                1e-14 (the reference's np.dot goes to BLAS,
 *                                whose summation order is not pinned).
 *
 * The RMM-DIIS restatement (rmmdiis.py) and the learned network
 * (unet.py) are pinned in their own headers.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif


/* ------------------------------------------------------------------ */
/* Halo fill.  Follows c/bc.c:17-135 (bc_init) and c/bc.c:138-250
 * (bc_unpack1) for a single rank: zero the padded buffer, paste the
 * interior at (lo0, lo1, lo2), then per axis i = 0, 1, 2 copy periodic
 * images.  For a periodic axis the right pad receives the first hi[i]
 * interior planes and the left pad the last lo[i] interior planes.  The
 * extent in the axes already processed is the full padded range when
 * `corners` is set (cfd == 0), the interior range otherwise.           */
void orc_halo_fill(const double* a, const long n[3], const long lo[3],
                   const long hi[3], const int pbc[3], int corners,
                   double* b)
{
    long m[3];
    for (int c = 0; c < 3; c++) m[c] = n[c] + lo[c] + hi[c];
    memset(b, 0, sizeof(double) * m[0] * m[1] * m[2]);
    for (long i0 = 0; i0 < n[0]; i0++)
        for (long i1 = 0; i1 < n[1]; i1++)
            memcpy(b + ((i0 + lo[0]) * m[1] + i1 + lo[1]) * m[2] + lo[2],
                   a + (i0 * n[1] + i1) * n[2], sizeof(double) * n[2]);

    long beg[3] = {lo[0], lo[1], lo[2]};
    long end[3] = {lo[0] + n[0], lo[1] + n[1], lo[2] + n[2]};
    for (int i = 0; i < 3; i++) {
        if (pbc[i]) {
            /* destination range along axis i and its source range */
            for (int side = 0; side < 2; side++) {
                long d0, d1, s0;
                if (side == 0) {          /* right pad <- first planes */
                    d0 = lo[i] + n[i];
                    d1 = d0 + hi[i];
                    s0 = lo[i];
                } else {                  /* left pad <- last planes */
                    d0 = 0;
                    d1 = lo[i];
                    s0 = lo[i] + n[i] - lo[i];
                }
                long r0[3] = {beg[0], beg[1], beg[2]};
                long r1[3] = {end[0], end[1], end[2]};
                r0[i] = d0;
                r1[i] = d1;
                for (long p0 = r0[0]; p0 < r1[0]; p0++)
                    for (long p1 = r0[1]; p1 < r1[1]; p1++)
                        for (long p2 = r0[2]; p2 < r1[2]; p2++) {
                            long q[3] = {p0, p1, p2};
                            q[i] = s0 + (q[i] - d0);
                            b[(p0 * m[1] + p1) * m[2] + p2] =
                                b[(q[0] * m[1] + q[1]) * m[2] + q[2]];
                        }
            }
        }
        if (corners) {
            beg[i] = 0;
            end[i] = m[i];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Complex halo fill with Bloch phases.  As orc_halo_fill, on interleaved
 * (re, im) values, with the periodic copies multiplied by the phase of
 * their direction (c/bc.c:227-247 with c/bmgs/translate.c:27-47,
 * translatemz): the right pad of axis i receives phase_cd[i][0] times the
 * first planes, the left pad phase_cd[i][1] times the last planes.  The
 * product is the C99 complex product of the reference's build:
 * re = pr*sr - pi*si, im = pr*si + pi*sr, two roundings each.  With
 * `corners`, axis i copies the pads of the axes before it, so a corner
 * value carries the phases of both axes, axis 0's applied first.
 * phase: 12 doubles, [axis][direction][re, im].                         */
void orc_halo_fill_z(const double* a, const long n[3], const long lo[3],
                     const long hi[3], const int pbc[3], int corners,
                     const double* phase, double* b)
{
    long m[3];
    for (int c = 0; c < 3; c++) m[c] = n[c] + lo[c] + hi[c];
    memset(b, 0, sizeof(double) * 2 * m[0] * m[1] * m[2]);
    for (long i0 = 0; i0 < n[0]; i0++)
        for (long i1 = 0; i1 < n[1]; i1++)
            memcpy(b + 2 * (((i0 + lo[0]) * m[1] + i1 + lo[1]) * m[2] + lo[2]),
                   a + 2 * ((i0 * n[1] + i1) * n[2]), sizeof(double) * 2 * n[2]);

    long beg[3] = {lo[0], lo[1], lo[2]};
    long end[3] = {lo[0] + n[0], lo[1] + n[1], lo[2] + n[2]};
    for (int i = 0; i < 3; i++) {
        if (pbc[i]) {
            for (int side = 0; side < 2; side++) {
                long d0, d1, s0;
                if (side == 0) {          /* right pad <- first planes */
                    d0 = lo[i] + n[i];
                    d1 = d0 + hi[i];
                    s0 = lo[i];
                } else {                  /* left pad <- last planes */
                    d0 = 0;
                    d1 = lo[i];
                    s0 = lo[i] + n[i] - lo[i];
                }
                const double pr = phase[4 * i + 2 * side];
                const double pim = phase[4 * i + 2 * side + 1];
                long r0[3] = {beg[0], beg[1], beg[2]};
                long r1[3] = {end[0], end[1], end[2]};
                r0[i] = d0;
                r1[i] = d1;
                for (long p0 = r0[0]; p0 < r1[0]; p0++)
                    for (long p1 = r0[1]; p1 < r1[1]; p1++)
                        for (long p2 = r0[2]; p2 < r1[2]; p2++) {
                            long q[3] = {p0, p1, p2};
                            q[i] = s0 + (q[i] - d0);
                            const double* src = b + 2 * ((q[0] * m[1] + q[1]) * m[2] + q[2]);
                            double* dst = b + 2 * ((p0 * m[1] + p1) * m[2] + p2);
                            const double sr = src[0], si = src[1];
                            const double re1 = pr * sr, re2 = pim * si;
                            const double im1 = pr * si, im2 = pim * sr;
                            dst[0] = re1 - re2;
                            dst[1] = im1 + im2;
                        }
            }
        }
        if (corners) {
            beg[i] = 0;
            end[i] = m[i];
        }
    }
}

/* Complex FD stencil (c/bmgs/fd.c:16-87 built with T = double_complex,
 * c/operators.c:109-157 with phases): per band the phased halo, then per
 * point x += a[p + off[c]] * coefs[c] from 0 in coefficient order -- a
 * real coefficient times a complex value is (re * c, im * c), so each
 * part is summed exactly as the real stencil sums it.  in/out hold
 * interleaved (re, im) values.                                         */
void orc_fdz(int ncoefs, const double* coefs, const long* off_c,
             const long n[3], const int pbc[3], const double* phase, long nb,
             const double* in, double* out)
{
    long r = 0;
    int corners = 0;
    for (int c = 0; c < ncoefs; c++) {
        int nonzero = 0;
        for (int k = 0; k < 3; k++) {
            long d = off_c[3 * c + k];
            if (d > r) r = d;
            if (d != 0) nonzero++;
        }
        if (nonzero >= 2) corners = 1;
    }
    long lo[3] = {r, r, r}, hi[3] = {r, r, r};
    long m[3] = {n[0] + 2 * r, n[1] + 2 * r, n[2] + 2 * r};
    long ng = n[0] * n[1] * n[2];
    long* flat = (long*)malloc(sizeof(long) * ncoefs);
    for (int c = 0; c < ncoefs; c++)
        flat[c] = (off_c[3 * c] * m[1] + off_c[3 * c + 1]) * m[2] +
                  off_c[3 * c + 2];
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * m[0] * m[1] * m[2]);
#pragma omp for schedule(static)
        for (long x = 0; x < nb; x++) {
            orc_halo_fill_z(in + 2 * x * ng, n, lo, hi, pbc, corners, phase, buf);
            double* o = out + 2 * x * ng;
            for (long i0 = 0; i0 < n[0]; i0++)
                for (long i1 = 0; i1 < n[1]; i1++) {
                    const double* a =
                        buf + 2 * (((i0 + r) * m[1] + i1 + r) * m[2] + r);
                    for (long i2 = 0; i2 < n[2]; i2++) {
                        double sr = 0.0, si = 0.0;
                        for (int c = 0; c < ncoefs; c++) {
                            const double* v = a + 2 * (i2 + flat[c]);
                            sr += v[0] * coefs[c];
                            si += v[1] * coefs[c];
                        }
                        *o++ = sr;
                        *o++ = si;
                    }
                }
        }
        free(buf);
    }
    free(flat);
}

/* ------------------------------------------------------------------ */
/* FD stencil.  Follows c/bmgs/fd.c:16-87 and c/operators.c:109-157:
 * per band, fill the halo (padding r on every side, corners only when the
 * stencil has off-axis points), then b[p] = sum_c coefs[c] * a[p + off[c]],
 * accumulated from 0.0 in coefficient order.  off_c are (d0, d1, d2)
 * triples.                                                             */
void orc_fd(int ncoefs, const double* coefs, const long* off_c,
            const long n[3], const int pbc[3], long nb, const double* in,
            double* out)
{
    long r = 0;
    int corners = 0;
    for (int c = 0; c < ncoefs; c++) {
        int nonzero = 0;
        for (int k = 0; k < 3; k++) {
            long d = off_c[3 * c + k];
            if (d > r) r = d;
            if (d != 0) nonzero++;
        }
        if (nonzero >= 2) corners = 1;
    }
    long lo[3] = {r, r, r}, hi[3] = {r, r, r};
    long m[3] = {n[0] + 2 * r, n[1] + 2 * r, n[2] + 2 * r};
    long ng = n[0] * n[1] * n[2];
    long* flat = (long*)malloc(sizeof(long) * ncoefs);
    for (int c = 0; c < ncoefs; c++)
        flat[c] = (off_c[3 * c] * m[1] + off_c[3 * c + 1]) * m[2] +
                  off_c[3 * c + 2];
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * m[0] * m[1] * m[2]);
#pragma omp for schedule(static)
        for (long x = 0; x < nb; x++) {
            orc_halo_fill(in + x * ng, n, lo, hi, pbc, corners, buf);
            double* o = out + x * ng;
            for (long i0 = 0; i0 < n[0]; i0++)
                for (long i1 = 0; i1 < n[1]; i1++) {
                    const double* a =
                        buf + ((i0 + r) * m[1] + i1 + r) * m[2] + r;
                    for (long i2 = 0; i2 < n[2]; i2++) {
                        double s = 0.0;
                        for (int c = 0; c < ncoefs; c++)
                            s += a[i2 + flat[c]] * coefs[c];
                        *o++ = s;
                    }
                }
        }
        free(buf);
    }
    free(flat);
}

/* ------------------------------------------------------------------ */
/* Operator.relax (c/operators.c:49-91) with bmgs_relax (c/bmgs/relax.c:
 * 7-73).  Every sweep fills the halo from fun (c/bc.c, as orc_halo_fill),
 * then walks the interior in lexicographic order with the stencil minus
 * its centre (coefficient 0), summed from 0.0 in coefficient order:
 *   method 1 (Gauss-Seidel): x = (src - x) * (1 / coef_0), written to fun
 *     and into the padded buffer, so later points see it;
 *   method 2 (weighted Jacobi): fun = (1 - w) fun + w (src - x) / coef_0,
 *     all neighbours from the buffer of the previous sweep.            */
void orc_relax(int method, int ncoefs, const double* coefs, const long* off_c,
               const long n[3], const int pbc[3], long nb, double* fun,
               const double* src, int nrelax, double w)
{
    long r = 0;
    int corners = 0;
    for (int c = 0; c < ncoefs; c++) {
        int nonzero = 0;
        for (int k = 0; k < 3; k++) {
            long d = off_c[3 * c + k];
            if (d > r) r = d;
            if (d != 0) nonzero++;
        }
        if (nonzero >= 2) corners = 1;
    }
    long lo[3] = {r, r, r}, hi[3] = {r, r, r};
    long m[3] = {n[0] + 2 * r, n[1] + 2 * r, n[2] + 2 * r};
    long ng = n[0] * n[1] * n[2];
    long* flat = (long*)malloc(sizeof(long) * ncoefs);
    for (int c = 0; c < ncoefs; c++)
        flat[c] = (off_c[3 * c] * m[1] + off_c[3 * c + 1]) * m[2] +
                  off_c[3 * c + 2];
    const double coef = 1.0 / coefs[0];
    double* buf = (double*)malloc(sizeof(double) * m[0] * m[1] * m[2]);
    for (long x = 0; x < nb; x++) {
        double* b = fun + x * ng;
        const double* sr = src + x * ng;
        for (int it = 0; it < nrelax; it++) {
            orc_halo_fill(b, n, lo, hi, pbc, corners, buf);
            long p = 0;
            for (long i0 = 0; i0 < n[0]; i0++)
                for (long i1 = 0; i1 < n[1]; i1++) {
                    double* a = buf + ((i0 + r) * m[1] + i1 + r) * m[2] + r;
                    for (long i2 = 0; i2 < n[2]; i2++, p++) {
                        double s = 0.0;
                        for (int c = 1; c < ncoefs; c++)
                            s += a[i2 + flat[c]] * coefs[c];
                        if (method == 1) {
                            s = (sr[p] - s) * coef;
                            b[p] = s;
                            a[i2] = s;
                        } else {
                            b[p] = (1.0 - w) * b[p] + w * (sr[p] - s) / coefs[0];
                        }
                    }
                }
        }
    }
    free(buf);
    free(flat);
}

/* ------------------------------------------------------------------ */
/* Restriction.  Follows c/bmgs/restrict.c:18-97,137-154: three 1-D
 * passes (last axis first), each writing its result with the processed
 * axis moved to the front, so that after three passes the layout is back
 * in C order.  Weights are the Lagrange weights of restrict.c:37-57.   */
static double rst1(int k, const double* a)
{
    if (k == 2)
        return 0.5 * (a[0] + 0.5 * (a[1] + a[-1]));
    if (k == 4)
        return 0.5 * (a[0] + 0.5625 * (a[1] + a[-1]) +
                      -0.0625 * (a[3] + a[-3]));
    if (k == 6)
        return 0.5 * (a[0] + 0.58593750 * (a[1] + a[-1]) +
                      -0.09765625 * (a[3] + a[-3]) +
                      0.01171875 * (a[5] + a[-5]));
    return 0.5 * (a[0] + 0.59814453125 * (a[1] + a[-1]) +
                  -0.11962890625 * (a[3] + a[-3]) +
                  0.02392578125 * (a[5] + a[-5]) +
                  -0.00244140625 * (a[7] + a[-7]));
}

/* one pass: `lines` lines of length len_in; nout outputs per line;
 * output transposed: b[j + lines * i]                                  */
static void rst_pass(int k, const double* a, long len_in, long nout,
                     long lines, double* b)
{
    for (long j = 0; j < lines; j++) {
        const double* aa = a + j * len_in + k - 1;
        for (long i = 0; i < nout; i++)
            b[j + lines * i] = rst1(k, aa + 2 * i);
    }
}

static void ip_pass(int k, const double* a, long nin, long lines,
                    const int skip[2], double* b)
{
    long len_in = k - 1 - skip[1] + nin;
    for (long j = 0; j < lines; j++) {
        const double* aa = a + j * len_in + k / 2 - 1;
        long o = -skip[0];          /* output position of the even point */
        for (long i = 0; i < nin; i++, aa++, o += 2) {
            if (!(i == 0 && skip[0]))
                b[j + lines * o] = aa[0];
            if (!(i == nin - 1 && skip[1])) {
                double v;
                if (k == 2)
                    v = 0.5 * (aa[0] + aa[1]);
                else if (k == 4)
                    v = (0.5625 * (aa[0] + aa[1]) +
                         -0.0625 * (aa[-1] + aa[2]));
                else if (k == 6)
                    v = (0.58593750 * (aa[0] + aa[1]) +
                         -0.09765625 * (aa[-1] + aa[2]) +
                         0.01171875 * (aa[-2] + aa[3]));
                else
                    v = (0.59814453125 * (aa[0] + aa[1]) +
                         -0.11962890625 * (aa[-1] + aa[2]) +
                         0.02392578125 * (aa[-2] + aa[3]) +
                         -0.00244140625 * (aa[-3] + aa[4]));
                b[j + lines * (o + 1)] = v;
            }
        }
    }
}

/* Transformer apply (c/transformers.c:54-123): per band halo fill with
 * corners (cfd == 0) using pads lo/hi, then restrict or interpolate.
 * skip is [3][2] (interpolation only).                                 */
void orc_transform(int k, int interpolate, const long n[3], const long lo[3],
                   const long hi[3], const int pbc[3], const int* skip,
                   const long nout[3], long nb, const double* in,
                   double* out)
{
    long m[3] = {n[0] + lo[0] + hi[0], n[1] + lo[1] + hi[1],
                 n[2] + lo[2] + hi[2]};
    long ng = n[0] * n[1] * n[2];
    long ngo = nout[0] * nout[1] * nout[2];
    long ng2 = m[0] * m[1] * m[2];
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * ng2);
        double* w = (double*)malloc(sizeof(double) * ng2 * 16);
        double* t = (double*)malloc(sizeof(double) * ng2 * 16);
#pragma omp for schedule(static)
        for (long x = 0; x < nb; x++) {
            orc_halo_fill(in + x * ng, n, lo, hi, pbc, 1, buf);
            double* o = out + x * ngo;
            if (!interpolate) {
                long e = 2 * k - 3;
                long o2 = (m[2] - e) / 2, o1 = (m[1] - e) / 2,
                     o0 = (m[0] - e) / 2;
                rst_pass(k, buf, m[2], o2, m[0] * m[1], w);
                rst_pass(k, w, m[1], o1, m[0] * o2, t);
                rst_pass(k, t, m[0], o0, o1 * o2, o);
            } else {
                long e = k - 1;
                const int* s2 = skip + 4;
                const int* s1 = skip + 2;
                const int* s0 = skip;
                long f2 = (m[2] - e) * 2 - s2[0] + s2[1];
                long f1 = (m[1] - e) * 2 - s1[0] + s1[1];
                ip_pass(k, buf, m[2] - e + s2[1], m[0] * m[1], s2, t);
                ip_pass(k, t, m[1] - e + s1[1], m[0] * f2, s1, w);
                ip_pass(k, w, m[0] - e + s0[1], f1 * f2, s0, o);
            }
        }
        free(t);
        free(w);
        free(buf);
    }
}

/* Complex Transformer apply with Bloch phases (c/transformers.c:54-123
 * with bmgs_restrictz / bmgs_interpolatez): the phased halo of
 * orc_halo_fill_z, then the same three passes; a real weight times a
 * complex value and complex sums are per part, so the passes run on the
 * real and the imaginary parts separately.  in/out interleaved.         */
void orc_transform_z(int k, int interpolate, const long n[3], const long lo[3],
                     const long hi[3], const int pbc[3], const int* skip,
                     const long nout[3], const double* phase, long nb,
                     const double* in, double* out)
{
    long m[3] = {n[0] + lo[0] + hi[0], n[1] + lo[1] + hi[1],
                 n[2] + lo[2] + hi[2]};
    long ng = n[0] * n[1] * n[2];
    long ngo = nout[0] * nout[1] * nout[2];
    long ng2 = m[0] * m[1] * m[2];
#pragma omp parallel
    {
        double* zbuf = (double*)malloc(sizeof(double) * 2 * ng2);
        double* buf = (double*)malloc(sizeof(double) * ng2);
        double* w = (double*)malloc(sizeof(double) * ng2 * 16);
        double* t = (double*)malloc(sizeof(double) * ng2 * 16);
        double* o = (double*)malloc(sizeof(double) * ngo);
#pragma omp for schedule(static)
        for (long x = 0; x < nb; x++) {
            orc_halo_fill_z(in + 2 * x * ng, n, lo, hi, pbc, 1, phase, zbuf);
            for (int part = 0; part < 2; part++) {
                for (long g = 0; g < ng2; g++) buf[g] = zbuf[2 * g + part];
                if (!interpolate) {
                    long e = 2 * k - 3;
                    long o2 = (m[2] - e) / 2, o1 = (m[1] - e) / 2,
                         o0 = (m[0] - e) / 2;
                    rst_pass(k, buf, m[2], o2, m[0] * m[1], w);
                    rst_pass(k, w, m[1], o1, m[0] * o2, t);
                    rst_pass(k, t, m[0], o0, o1 * o2, o);
                } else {
                    long e = k - 1;
                    const int* s2 = skip + 4;
                    const int* s1 = skip + 2;
                    const int* s0 = skip;
                    long f2 = (m[2] - e) * 2 - s2[0] + s2[1];
                    long f1 = (m[1] - e) * 2 - s1[0] + s1[1];
                    ip_pass(k, buf, m[2] - e + s2[1], m[0] * m[1], s2, t);
                    ip_pass(k, t, m[1] - e + s1[1], m[0] * f2, s1, w);
                    ip_pass(k, w, m[0] - e + s0[1], f1 * f2, s0, o);
                }
                double* dst = out + 2 * x * ngo;
                for (long g = 0; g < ngo; g++) dst[2 * g + part] = o[g];
            }
        }
        free(o);
        free(t);
        free(w);
        free(buf);
        free(zbuf);
    }
}

/* ------------------------------------------------------------------ */
/* Cubic radial spline.  Follows c/bmgs/spline.c:11-58 and
 * c/spline.c:121-135: nbins = len(f) - 1, dr = rcut / nbins, natural
 * spline with zero slope at r = 0 (f2[0] = -0.5 start of the recursion),
 * four polynomial coefficients per bin, and a zero bin past the end.
 * data has 4 * (nbins + 1) doubles.                                    */
void orc_spline_build(double rcut, int npts, const double* f, double* data)
{
    int nbins = npts - 1;
    double dr = rcut / nbins;
    double c = 3.0 / (dr * dr);
    double* f2 = (double*)malloc(sizeof(double) * (nbins + 1));
    double* u = (double*)malloc(sizeof(double) * nbins);
    f2[0] = -0.5;
    u[0] = (f[1] - f[0]) * c;
    for (int b = 1; b < nbins; b++) {
        double p = 0.5 * f2[b - 1] + 2.0;
        f2[b] = -0.5 / p;
        u[b] = ((f[b + 1] - 2.0 * f[b] + f[b - 1]) * c - 0.5 * u[b - 1]) / p;
    }
    f2[nbins] = ((f[nbins - 1] * c - 0.5 * u[nbins - 1]) /
                 (0.5 * f2[nbins - 1] + 1.0));
    for (int b = nbins - 1; b >= 0; b--) f2[b] = f2[b] * f2[b + 1] + u[b];
    for (int b = 0; b < nbins; b++) {
        data[4 * b + 0] = f[b];
        data[4 * b + 1] =
            (f[b + 1] - f[b]) / dr - (f2[b] / 3 + f2[b + 1] / 6) * dr;
        data[4 * b + 2] = 0.5 * f2[b];
        data[4 * b + 3] = (f2[b + 1] - f2[b]) / (6 * dr);
    }
    for (int q = 0; q < 4; q++) data[4 * nbins + q] = 0.0;
    free(u);
    free(f2);
}

double orc_spline_value(const double* data, int nbins, double dr, double r)
{
    int b = r / dr;
    if (b >= nbins) return 0.0;
    double u = r - b * dr;
    const double* s = data + 4 * b;
    return s[0] + u * (s[1] + u * (s[2] + u * s[3]));
}

/* Real solid spherical harmonics times f, l = 0..4, m ordered as in
 * c/bmgs/spherical_harmonics.h:5-48.  The normalisation constants are the
 * standard real-harmonic prefactors written as decimal literals.       */
static void solid_harmonics(int l, double f, double x, double y, double z,
                            double r2, double* p)
{
    switch (l) {
    case 0:
        p[0] = f * 0.28209479177387814;
        break;
    case 1:
        p[0] = f * 0.48860251190291992 * y;
        p[1] = f * 0.48860251190291992 * z;
        p[2] = f * 0.48860251190291992 * x;
        break;
    case 2:
        p[0] = f * 1.0925484305920792 * x * y;
        p[1] = f * 1.0925484305920792 * y * z;
        p[2] = f * 0.31539156525252005 * (-r2 + 3 * z * z);
        p[3] = f * 1.0925484305920792 * x * z;
        p[4] = f * 0.54627421529603959 * (-y * y + x * x);
        break;
    case 3:
        p[0] = f * 0.59004358992664352 * (-y * y * y + 3 * x * x * y);
        p[1] = f * 2.8906114426405538 * x * y * z;
        p[2] = f * 0.45704579946446577 * (5 * y * z * z - y * r2);
        p[3] = f * 0.3731763325901154 * (-3 * z * r2 + 5 * z * z * z);
        p[4] = f * 0.45704579946446577 * (-x * r2 + 5 * x * z * z);
        p[5] = f * 1.4453057213202769 * (-y * y * z + x * x * z);
        p[6] = f * 0.59004358992664352 * (x * x * x - 3 * x * y * y);
        break;
    default:
        p[0] = f * 2.5033429417967046 * (x * x * x * y - x * y * y * y);
        p[1] = f * 1.7701307697799307 * (3 * x * x * y * z - y * y * y * z);
        p[2] = f * 0.94617469575756008 * (-x * y * r2 + 7 * x * y * z * z);
        p[3] = f * 0.66904654355728921 * (-3 * y * z * r2 + 7 * y * z * z * z);
        p[4] = f * 0.10578554691520431 *
               (3 * r2 * r2 - 30 * z * z * r2 + 35 * z * z * z * z);
        p[5] = f * 0.66904654355728921 * (7 * x * z * z * z - 3 * x * z * r2);
        p[6] = f * 0.47308734787878004 *
               (y * y * r2 + 7 * x * x * z * z - x * x * r2 - 7 * y * y * z * z);
        p[7] = f * 1.7701307697799307 * (x * x * x * z - 3 * x * y * y * z);
        p[8] = f * 0.62583573544917614 *
               (-6 * x * x * y * y + x * x * x * x + y * y * y * y);
        break;
    }
}

/* spline_to_grid (c/lfc.c:785-875): tabulate A_gm over the box
 * [beg, end) for points with r < rcut; return the z-runs in G_B as
 * (start, stop) pairs of flat indices of the local grid (corner
 * gdcorner, size n).  A_gm and G_B must be sized for the full box.
 * Returns the number of rows of A_gm; *nB receives len(G_B).           */
long orc_spline_to_grid(const double* data, int nbins, double dr, int l,
                        const long beg[3], const long end[3],
                        const double pos[3], const double h_cv[9],
                        const long n[3], const long corner[3], double* A_gm,
                        int* G_B, long* nB_out)
{
    int nm = 2 * l + 1;
    double rcut = dr * nbins;
    long ngm = 0, nB = 0;
    long G = -corner[2] + n[2] * (beg[1] - corner[1] + n[1] * (beg[0] - corner[0]));
    for (long g0 = beg[0]; g0 < end[0]; g0++) {
        for (long g1 = beg[1]; g1 < end[1]; g1++) {
            long first = -1, last = -1;
            for (long g2 = beg[2]; g2 < end[2]; g2++) {
                double x = h_cv[0] * g0 + h_cv[3] * g1 + h_cv[6] * g2 - pos[0];
                double y = h_cv[1] * g0 + h_cv[4] * g1 + h_cv[7] * g2 - pos[1];
                double z = h_cv[2] * g0 + h_cv[5] * g1 + h_cv[8] * g2 - pos[2];
                double r2 = x * x + y * y + z * z;
                double r = sqrt(r2);
                if (r < rcut) {
                    if (first < 0) first = g2;
                    last = g2;
                    double f = orc_spline_value(data, nbins, dr, r);
                    solid_harmonics(l, f, x, y, z, r2, A_gm + ngm);
                    ngm += nm;
                }
            }
            if (last >= 0) {
                G_B[nB++] = (int)(G + first);
                G_B[nB++] = (int)(G + last + 1);
            }
            G += n[2];
        }
        G += n[2] * (n[1] - end[1] + beg[1]);
    }
    *nB_out = nB;
    return ngm / nm;
}

/* ------------------------------------------------------------------ */
/* LFC boundary sweep (c/lfc.h:38-87).  Volumes W have values A_W (rows
 * of nm_W[W] functions, concatenated at offset A_off[W]) and first
 * function index M_W[W].  The sorted boundary list (G_B, W_B) holds
 * entering (W >= 0) and leaving (-W-1) events.  Between two events the
 * set of active volumes is constant; a volume's row pointer advances by
 * the number of points it has covered.                                  */
typedef struct {
    long pos;   /* current row offset into the concatenated values */
    int nm;
    int M;
    int W;
} active_vol;

typedef void (*segment_fn)(void* ctx, long Ga, long Gb, const active_vol* v,
                           int ni);

static void lfc_sweep(int nW, const long* A_off, const int* nm_W,
                      const int* M_W, long nB, const int* G_B,
                      const int* W_B, segment_fn fn, void* ctx)
{
    active_vol* vol_W = (active_vol*)malloc(sizeof(active_vol) * (nW + 1));
    active_vol* vol_i = (active_vol*)malloc(sizeof(active_vol) * (nW + 1));
    int* i_W = (int*)malloc(sizeof(int) * (nW + 1));
    for (int W = 0; W < nW; W++) {
        vol_W[W].pos = A_off[W];
        vol_W[W].nm = nm_W[W];
        vol_W[W].M = M_W[W];
        vol_W[W].W = W;
    }
    long Ga = 0;
    int ni = 0;
    for (long B = 0; B < nB; B++) {
        long Gb = G_B[B];
        long nG = Gb - Ga;
        if (nG > 0) {
            fn(ctx, Ga, Gb, vol_i, ni);
            for (int i = 0; i < ni; i++) vol_i[i].pos += nG * vol_i[i].nm;
        }
        int Wn = W_B[B];
        if (Wn >= 0) {
            vol_i[ni] = vol_W[Wn];
            i_W[Wn] = ni;
            ni++;
        } else {
            int Wold = -1 - Wn;
            int iold = i_W[Wold];
            vol_W[Wold].pos = vol_i[iold].pos;
            ni--;
            vol_i[iold] = vol_i[ni];
            i_W[vol_i[iold].W] = iold;
        }
        Ga = Gb;
    }
    free(i_W);
    free(vol_i);
    free(vol_W);
}

typedef struct {
    const double* A;
    const double* a_G;
    double* c_M;
    double dv;
    const double* cc_M;
    double* aa_G;
} lfc_ctx;

static void integrate_seg(void* p, long Ga, long Gb, const active_vol* v,
                          int ni)
{
    lfc_ctx* ctx = (lfc_ctx*)p;
    for (int i = 0; i < ni; i++) {
        const double* A = ctx->A + v[i].pos;
        double* c = ctx->c_M + v[i].M;
        int nm = v[i].nm;
        long gm = 0;
        for (long G = Ga; G < Gb; G++) {
            double av = ctx->a_G[G] * ctx->dv;
            for (int m = 0; m < nm; m++, gm++) c[m] += av * A[gm];
        }
    }
}

static void add_seg(void* p, long Ga, long Gb, const active_vol* v, int ni)
{
    lfc_ctx* ctx = (lfc_ctx*)p;
    for (int i = 0; i < ni; i++) {
        const double* A = ctx->A + v[i].pos;
        const double* c = ctx->cc_M + v[i].M;
        int nm = v[i].nm;
        long gm = 0;
        for (long G = Ga; G < Gb; G++)
            for (int m = 0; m < nm; m++, gm++) ctx->aa_G[G] += A[gm] * c[m];
    }
}

/* c_xM += dv * sum_G a_xG[G] A(G, M)   (c/lfc.c:361-425, real case) */
void orc_lfc_integrate(int nW, const double* A, const long* A_off,
                       const int* nm_W, const int* M_W, long nB,
                       const int* G_B, const int* W_B, double dv, long nx,
                       long nG, const double* a_xG, long nM, double* c_xM)
{
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nx; x++) {
        lfc_ctx ctx = {A, a_xG + x * nG, c_xM + x * nM, dv, 0, 0};
        lfc_sweep(nW, A_off, nm_W, M_W, nB, G_B, W_B, integrate_seg, &ctx);
    }
}

/* a_xG[G] += sum_M A(G, M) c_xM   (c/lfc.c:724-783, real case) */
void orc_lfc_add(int nW, const double* A, const long* A_off,
                 const int* nm_W, const int* M_W, long nB, const int* G_B,
                 const int* W_B, long nx, long nG, const double* c_xM,
                 long nM, double* a_xG)
{
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nx; x++) {
        lfc_ctx ctx = {A, 0, 0, 0.0, c_xM + x * nM, a_xG + x * nG};
        lfc_sweep(nW, A_off, nm_W, M_W, nB, G_B, W_B, add_seg, &ctx);
    }
}

/* Complex (Bloch) functions (c/lfc.c:400-423, 758-781): volume W's
 * values carry the phase phase_W[W] (interleaved re, im) of the k-point.
 * integrate: vphase = phase * dv (each part times dv), per point
 * av = a * vphase (C99 complex product), c[m] += av * A (each part);
 * add: per point and volume t = sum_m A c[m] from 0, then
 * a += t * conj(phase).                                                */
typedef struct {
    const double* A;
    const double* a_G;
    double* c_M;
    double dv;
    const double* cc_M;
    double* aa_G;
    const double* phase_W;
} lfcz_ctx;

static void integrate_seg_z(void* p, long Ga, long Gb, const active_vol* v,
                            int ni)
{
    lfcz_ctx* ctx = (lfcz_ctx*)p;
    for (int i = 0; i < ni; i++) {
        const double* A = ctx->A + v[i].pos;
        double* c = ctx->c_M + 2 * v[i].M;
        int nm = v[i].nm;
        const double vr = ctx->phase_W[2 * v[i].W] * ctx->dv;
        const double vi = ctx->phase_W[2 * v[i].W + 1] * ctx->dv;
        long gm = 0;
        for (long G = Ga; G < Gb; G++) {
            const double ar = ctx->a_G[2 * G], ai = ctx->a_G[2 * G + 1];
            const double r1 = ar * vr, r2 = ai * vi;
            const double i1 = ar * vi, i2 = ai * vr;
            const double avr = r1 - r2, avi = i1 + i2;
            for (int m = 0; m < nm; m++, gm++) {
                c[2 * m] += avr * A[gm];
                c[2 * m + 1] += avi * A[gm];
            }
        }
    }
}

static void add_seg_z(void* p, long Ga, long Gb, const active_vol* v, int ni)
{
    lfcz_ctx* ctx = (lfcz_ctx*)p;
    for (int i = 0; i < ni; i++) {
        const double* A = ctx->A + v[i].pos;
        const double* c = ctx->cc_M + 2 * v[i].M;
        int nm = v[i].nm;
        const double pr = ctx->phase_W[2 * v[i].W];
        const double pim = -ctx->phase_W[2 * v[i].W + 1];
        long gm = 0;
        for (long G = Ga; G < Gb; G++) {
            double tr = 0.0, ti = 0.0;
            for (int m = 0; m < nm; m++, gm++) {
                tr += A[gm] * c[2 * m];
                ti += A[gm] * c[2 * m + 1];
            }
            const double r1 = tr * pr, r2 = ti * pim;
            const double i1 = tr * pim, i2 = ti * pr;
            ctx->aa_G[2 * G] += r1 - r2;
            ctx->aa_G[2 * G + 1] += i1 + i2;
        }
    }
}

void orc_lfc_integrate_z(int nW, const double* A, const long* A_off,
                         const int* nm_W, const int* M_W, long nB,
                         const int* G_B, const int* W_B, double dv, long nx,
                         long nG, const double* a_xG, long nM, double* c_xM,
                         const double* phase_W)
{
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nx; x++) {
        lfcz_ctx ctx = {A, a_xG + 2 * x * nG, c_xM + 2 * x * nM, dv, 0, 0,
                        phase_W};
        lfc_sweep(nW, A_off, nm_W, M_W, nB, G_B, W_B, integrate_seg_z, &ctx);
    }
}

void orc_lfc_add_z(int nW, const double* A, const long* A_off,
                   const int* nm_W, const int* M_W, long nB, const int* G_B,
                   const int* W_B, long nx, long nG, const double* c_xM,
                   long nM, double* a_xG, const double* phase_W)
{
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nx; x++) {
        lfcz_ctx ctx = {A, 0, 0, 0.0, c_xM + 2 * x * nM, a_xG + 2 * x * nG,
                        phase_W};
        lfc_sweep(nW, A_off, nm_W, M_W, nB, G_B, W_B, add_seg_z, &ctx);
    }
}

/* ------------------------------------------------------------------ */
/* Local potential: Htpsit_G += psit_G * vt_G, band by band
 * (gpaw/hamiltonian.py:397-424).                                        */
void orc_add_local_potential(long nb, long nG, const double* psi,
                             const double* v, double* hpsi)
{
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nb; x++) {
        const double* p = psi + x * nG;
        double* h = hpsi + x * nG;
        for (long G = 0; G < nG; G++) {
            double t = p[G] * v[G];
            h[G] = h[G] + t;
        }
    }
}

/* Block-diagonal projector coefficients c[x, M_a + j] =
 * sum_i P[x, M_a + i] dH_a[i, j], summed in i order with two roundings.
 * (np.dot in gpaw/hamiltonian.py:456-458; the reference delegates to
 * BLAS, so the summation order there is the BLAS library's.)           */
void orc_projector_coefs(long nb, long nM, int natoms, const int* M_a,
                         const int* ni_a, const long* dH_off,
                         const double* dH, const double* P, double* c)
{
    for (long x = 0; x < nb; x++)
        for (int a = 0; a < natoms; a++) {
            int ni = ni_a[a];
            const double* Pa = P + x * nM + M_a[a];
            const double* H = dH + dH_off[a];
            double* ca = c + x * nM + M_a[a];
            for (int j = 0; j < ni; j++) {
                double s = 0.0;
                for (int i = 0; i < ni; i++) s += Pa[i] * H[i * ni + j];
                ca[j] = s;
            }
        }
}

/* ------------------------------------------------------------------ */
/* Multigrid preconditioner (gpaw/preconditioner.py:34-60).  The caller
 * supplies the three levels' stencils and transformer set-ups; the
 * sequence of array operations and their rounding is the reference's:
 * numpy scalar factors are formed first (4*step, 16*step), then applied
 * elementwise.                                                          */
typedef struct {
    int ncoefs;
    const double* coefs;
    const long* off_c;
    long n[3];
} orc_stencil;

typedef struct {
    int k;
    int interpolate;
    long n[3];
    long lo[3];
    long hi[3];
    int skip[6];
    long nout[3];
} orc_transformer;

void orc_precondition(const orc_stencil* kin0, const orc_stencil* kin1,
                      const orc_stencil* kin2, const orc_transformer* rst0,
                      const orc_transformer* rst1,
                      const orc_transformer* ip2, const orc_transformer* ip1,
                      const int pbc[3], double step, long nb,
                      const double* R, double* out)
{
    long g0 = kin0->n[0] * kin0->n[1] * kin0->n[2];
    long g1 = kin1->n[0] * kin1->n[1] * kin1->n[2];
    long g2 = kin2->n[0] * kin2->n[1] * kin2->n[2];
    double* mR = (double*)malloc(sizeof(double) * nb * g0);
    double* d0 = (double*)malloc(sizeof(double) * nb * g0);
    double* q0 = (double*)malloc(sizeof(double) * nb * g0);
    double* r1 = (double*)malloc(sizeof(double) * nb * g1);
    double* d1 = (double*)malloc(sizeof(double) * nb * g1);
    double* q1 = (double*)malloc(sizeof(double) * nb * g1);
    double* r2 = (double*)malloc(sizeof(double) * nb * g2);
    double* d2 = (double*)malloc(sizeof(double) * nb * g2);
    double* q2 = (double*)malloc(sizeof(double) * nb * g2);
    double s4 = 4 * step, s16 = 16 * step, ms = -step;
    long i;

    for (i = 0; i < nb * g0; i++) mR[i] = -R[i];
    orc_transform(rst0->k, 0, rst0->n, rst0->lo, rst0->hi, pbc, rst0->skip,
                  rst0->nout, nb, mR, r1);
    for (i = 0; i < nb * g1; i++) d1[i] = s4 * r1[i];
    orc_fd(kin1->ncoefs, kin1->coefs, kin1->off_c, kin1->n, pbc, nb, d1, q1);
    for (i = 0; i < nb * g1; i++) q1[i] = q1[i] - r1[i];
    orc_transform(rst1->k, 0, rst1->n, rst1->lo, rst1->hi, pbc, rst1->skip,
                  rst1->nout, nb, q1, r2);
    for (i = 0; i < nb * g2; i++) d2[i] = s16 * r2[i];
    orc_fd(kin2->ncoefs, kin2->coefs, kin2->off_c, kin2->n, pbc, nb, d2, q2);
    for (i = 0; i < nb * g2; i++) q2[i] = q2[i] - r2[i];
    for (i = 0; i < nb * g2; i++) d2[i] = d2[i] - s16 * q2[i];
    orc_transform(ip2->k, 1, ip2->n, ip2->lo, ip2->hi, pbc, ip2->skip,
                  ip2->nout, nb, d2, q1);
    for (i = 0; i < nb * g1; i++) d1[i] = d1[i] - q1[i];
    orc_fd(kin1->ncoefs, kin1->coefs, kin1->off_c, kin1->n, pbc, nb, d1, q1);
    for (i = 0; i < nb * g1; i++) q1[i] = q1[i] - r1[i];
    for (i = 0; i < nb * g1; i++) d1[i] = d1[i] - s4 * q1[i];
    for (i = 0; i < nb * g1; i++) d1[i] = -d1[i];
    orc_transform(ip1->k, 1, ip1->n, ip1->lo, ip1->hi, pbc, ip1->skip,
                  ip1->nout, nb, d1, d0);
    orc_fd(kin0->ncoefs, kin0->coefs, kin0->off_c, kin0->n, pbc, nb, d0, q0);
    for (i = 0; i < nb * g0; i++) q0[i] = q0[i] - R[i];
    for (i = 0; i < nb * g0; i++) out[i] = -(d0[i] + ms * q0[i]);

    free(q2); free(d2); free(r2); free(q1); free(d1); free(r1);
    free(q0); free(d0); free(mR);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/*
 * CPU oracle for the moment-image hot path. TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP kernels and the CPU baseline timed by bench.py. See
 * mray_oracle.h for the precision modes.
 *
 * Build: gcc -std=gnu11 -O2 -fopenmp -ffp-contract=off (no -march, no
 * -ffast-math) so that float expressions round exactly like the reference's
 * g++ -O2 build on x86-64 (SSE2, FLT_EVAL_METHOD == 0).
 *
 * Where the reference mixes float and double (M_PI literals, 0.5 literals,
 * 1.0 / x, unqualified cos/sin/exp/atan2 on floats, which resolve to the
 * double overloads), the same promotion is written out explicitly.
 */
#include "mray_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define F_PI ((float)M_PI)

/* ------------------------------------------------------------------------ */
/* helpers: cut::min / cut::max / clamp (src/cut/cuda_math.h:72-85,825-828)  */
/* ------------------------------------------------------------------------ */
static inline float fminr(float x, float y) { return (x < y) ? x : y; }
static inline float fmaxr(float x, float y) { return (x > y) ? x : y; }
static inline float fclampr(float v, float lo, float hi) { return fminr(fmaxr(v, lo), hi); }
static inline int iclampr(int v, int lo, int hi)
{
    int a = (v > lo) ? v : lo;
    return (a < hi) ? a : hi;
}
static inline double dminr(double x, double y) { return (x < y) ? x : y; }
static inline double dmaxr(double x, double y) { return (x > y) ? x : y; }
static inline double dclampr(double v, double lo, double hi) { return dminr(dmaxr(v, lo), hi); }

/* float_complex (src/mese/complex_algebra.h:34-158) */
typedef struct
{
    float x, y;
} cf;
static inline cf cf_make(float x, float y)
{
    cf r;
    r.x = x;
    r.y = y;
    return r;
}
static inline cf cf_add(cf a, cf b) { return cf_make(a.x + b.x, a.y + b.y); }
static inline cf cf_sub(cf a, cf b) { return cf_make(a.x - b.x, a.y - b.y); }
static inline cf cf_mul(cf a, cf b) { return cf_make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
static inline cf cf_rmul(float s, cf a) { return cf_make(s * a.x, s * a.y); }
static inline cf cf_conj(cf a) { return cf_make(a.x, -a.y); }
static inline float cf_abs2(cf a) { return a.x * a.x + a.y * a.y; }
static inline cf cf_div(cf n, cf d)
{
    float factor = 1.0f / cf_abs2(d);
    return cf_make((n.x * d.x + n.y * d.y) * factor, (-n.x * d.y + n.y * d.x) * factor);
}
static inline cf cf_divr(cf n, float d)
{
    float factor = 1.0f / d;
    return cf_make(n.x * factor, n.y * factor);
}
/* euler_exp: the reference's unqualified cos/sin on a float resolve to the
 * double overloads (src/mese/complex_algebra.h:155-158). */
static inline cf euler_exp(float x) { return cf_make((float)cos((double)x), (float)sin((double)x)); }

typedef struct
{
    double x, y;
} cd;
static inline cd cd_make(double x, double y)
{
    cd r;
    r.x = x;
    r.y = y;
    return r;
}
static inline cd cd_add(cd a, cd b) { return cd_make(a.x + b.x, a.y + b.y); }
static inline cd cd_sub(cd a, cd b) { return cd_make(a.x - b.x, a.y - b.y); }
static inline cd cd_mul(cd a, cd b) { return cd_make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
static inline cd cd_rmul(double s, cd a) { return cd_make(s * a.x, s * a.y); }
static inline cd cd_conj(cd a) { return cd_make(a.x, -a.y); }
static inline double cd_abs2(cd a) { return a.x * a.x + a.y * a.y; }

static int set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        return nthreads;
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* camera / rays                                                             */
/* ------------------------------------------------------------------------ */

/* get_eye_ray_persp / _ortho (src/cut/raytracing.h:39-65), Mat4f*Vec4f
 * (src/cut/matrix.h:658-664), normalize (src/cut/cuda_math.h:502-511). */
void orc_eye_ray(const mray_camera *cam, float ndcx, float ndcy, float org[3], float dir[3])
{
    if (cam->orthographic)
    {
        for (int i = 0; i < 3; ++i)
        {
            float a = cam->position[i] + ndcx * cam->right[i];
            org[i] = a + ndcy * cam->up[i];
            dir[i] = cam->dir[i];
        }
    }
    else
    {
        org[0] = cam->position[0];
        org[1] = cam->position[1];
        org[2] = cam->position[2];
        float x = ndcx * cam->aspect * cam->scale;
        float y = ndcy * cam->scale;
        const float *M = cam->transform;
        float v[3];
        for (int r = 0; r < 3; ++r)
            v[r] = M[r * 4 + 0] * x + M[r * 4 + 1] * y + M[r * 4 + 2] * -1.0f + M[r * 4 + 3] * 0.0f;
        float inv = 1 / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        dir[0] = v[0] * inv;
        dir[1] = v[1] * inv;
        dir[2] = v[2] * inv;
    }
}

/* Pixel -> NDC without half-pixel offset (src/moment_raymarcher.cpp:133,298). */
void orc_pixel_ray(const mray_camera *cam, int W, int H, int x, int y, float org[3], float dir[3])
{
    float ndcx = x / (float)W * 2.f - 1.f;
    float ndcy = y / (float)H * 2.f - 1.f;
    orc_eye_ray(cam, ndcx, ndcy, org, dir);
}

/* Slab test with NaN-dropping operand order (src/cut/raytracing.h:67-86). */
int orc_intersect_aabb(const float org[3], const float dir[3], const float bmin[3], const float bmax[3],
                       float *t_near, float *t_far)
{
    float tmin[3], tmax[3];
    for (int i = 0; i < 3; ++i)
    {
        float invR = 1.0f / dir[i];
        float tbot = invR * (bmin[i] - org[i]);
        float ttop = invR * (bmax[i] - org[i]);
        tmin[i] = fminr(ttop, tbot);
        tmax[i] = fmaxr(ttop, tbot);
    }
    float largest_tmin = fmaxr(fmaxr(tmin[0], tmin[1]), fmaxr(tmin[0], tmin[2]));
    float smallest_tmax = fminr(fminr(tmax[0], tmax[1]), fminr(tmax[0], tmax[2]));
    *t_near = largest_tmin;
    *t_far = smallest_tmax;
    return smallest_tmax > largest_tmin;
}

/* ------------------------------------------------------------------------ */
/* volume sampling (src/volume.h:59-123)                                     */
/* ------------------------------------------------------------------------ */
static inline float vol_get(const orc_volume *v, int x, int y, int z)
{
    x = iclampr(x, 0, v->nx - 1);
    y = iclampr(y, 0, v->ny - 1);
    z = iclampr(z, 0, v->nz - 1);
    size_t idx = (size_t)z * (size_t)(v->ny * v->nx) + (size_t)(y * v->nx + x);
    if (v->dtype == MRAY_DTYPE_U8)
        return (float)(((const uint8_t *)v->data)[idx] / 255.f);
    return ((const float *)v->data)[idx];
}

float orc_volume_sample(const orc_volume *v, const float p[3])
{
    if (v->dtype == MRAY_DTYPE_PARTICLES)
        return orc_sph_sample((const orc_sph *)v->data, p);
    float ip[3], a[3];
    int il[3];
    const int n[3] = {v->nx, v->ny, v->nz};
    for (int i = 0; i < 3; ++i)
    {
        float np = (p[i] - v->dmin[i]) / (v->dmax[i] - v->dmin[i]);
        ip[i] = np * (float)n[i];
        a[i] = ip[i] - floorf(ip[i]);
        il[i] = (int)floorf(ip[i]);
    }
    int xl = il[0], yl = il[1], zl = il[2];
    float lll = vol_get(v, xl, yl, zl);
    float lrl = vol_get(v, xl, yl + 1, zl);
    float rrl = vol_get(v, xl + 1, yl + 1, zl);
    float rll = vol_get(v, xl + 1, yl, zl);
    float llr = vol_get(v, xl, yl, zl + 1);
    float lrr = vol_get(v, xl, yl + 1, zl + 1);
    float rrr = vol_get(v, xl + 1, yl + 1, zl + 1);
    float rlr = vol_get(v, xl + 1, yl, zl + 1);
    float ax = a[0], ay = a[1], az = a[2];
    return (1.f - ax) * (1.f - ay) * (1.f - az) * lll + ax * (1.f - ay) * (1.f - az) * rll +
           (1.f - ax) * ay * (1.f - az) * lrl + ax * ay * (1.f - az) * rrl + (1.f - ax) * (1.f - ay) * az * llr +
           ax * (1.f - ay) * az * rlr + (1.f - ax) * ay * az * lrr + ax * ay * az * rrr;
}

double orc_volume_sample_f64(const orc_volume *v, const double p[3])
{
    if (v->dtype == MRAY_DTYPE_PARTICLES)
    {
        const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
        return orc_sph_sample((const orc_sph *)v->data, pf);
    }
    double a[3];
    int il[3];
    const int n[3] = {v->nx, v->ny, v->nz};
    for (int i = 0; i < 3; ++i)
    {
        double ip = (p[i] - v->dmin[i]) / ((double)v->dmax[i] - v->dmin[i]) * n[i];
        double fl = floor(ip);
        a[i] = ip - fl;
        il[i] = (int)fl;
    }
    int xl = il[0], yl = il[1], zl = il[2];
    double ax = a[0], ay = a[1], az = a[2];
    double c[2][2][2];
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                c[dz][dy][dx] = vol_get(v, xl + dx, yl + dy, zl + dz);
    double r = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                r += (dx ? ax : 1.0 - ax) * (dy ? ay : 1.0 - ay) * (dz ? az : 1.0 - az) * c[dz][dy][dx];
    return r;
}

/* map_density (src/parameters.h:31-34) */
static inline float map_density(float s, const mray_gen_params *p)
{
    return fclampr((s - p->data_min) / (p->data_max - p->data_min), 0.0f, 1.0f);
}

/* DensityBound::set (src/common.h:143-147): truncating 16-bit quantisation */
static inline uint32_t pack_bound(float l, float u)
{
    uint16_t lo = (uint16_t)(l * 65535.f);
    uint16_t hi = (uint16_t)(u * 65535.f);
    return (uint32_t)lo | ((uint32_t)hi << 16);
}

/* ------------------------------------------------------------------------ */
/* generate (src/raymarch_generate.h)                                        */
/* ------------------------------------------------------------------------ */

/* update_moments (src/raymarch_generate.h:35-58) */
static void update_moments(int m, float *trig, float p_next, float phase_step_size, float fs, float bs)
{
    float p_prev = p_next - phase_step_size;
    float gradient = (bs - fs) / phase_step_size;
    float yintercept = fs - gradient * p_prev;

    for (int j = 1; j < m; ++j)
    {
        float jSq = (float)j * (float)j;
        cf common = cf_make(gradient / jSq, yintercept / (float)j);

        cf a = cf_add(common, cf_make(0, gradient * (float)j * p_next / jSq));
        trig[j] += cf_mul(a, euler_exp((float)(-j) * p_next)).x;
        cf b = cf_add(common, cf_make(0, gradient * (float)j * p_prev / jSq));
        trig[j] -= cf_mul(b, euler_exp((float)(-j) * p_prev)).x;
    }
    trig[0] = (float)((double)trig[0] + (0.5 * gradient * (double)(p_next * p_next) + (double)(yintercept * p_next)));
    trig[0] = (float)((double)trig[0] - (0.5 * gradient * (double)(p_prev * p_prev) + (double)(yintercept * p_prev)));
}

static inline void sample_at(const orc_volume *v, const float org[3], const float dir[3], float t, float out[3])
{
    for (int i = 0; i < 3; ++i)
        out[i] = org[i] + t * dir[i];
}

/* generate_ray (src/raymarch_generate.h:69-148) */
uint32_t orc_generate_ray(const orc_volume *v, const float org[3], const float dir[3], float znear, float *moments,
                          int m, const mray_gen_params *gp, int32_t *n_samples)
{
    float trig[512];
    for (int j = 0; j < m; ++j)
        trig[j] = 0.0f;

    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, v->dmin, v->dmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    int ns = 0;

    if (hit && (t_far > 0))
    {
        float step = gp->step_size;
        float bound_lo = 1.f, bound_hi = 0.f;
        float t = t_near;
        float pi_trange_inv = (float)(M_PI / (double)(t_far - t_near));
        float p = (float)(-M_PI);
        float phase_step_size = (float)(M_PI * (double)(step / (t_far - t_near)));
        float pos[3];

        sample_at(v, org, dir, t, pos);
        float fs = map_density(orc_volume_sample(v, pos), gp);
        ++ns;
        bound_lo = fminr(bound_lo, fs);
        bound_hi = fmaxr(bound_hi, fs);
        float bs;
        while (t + step < t_far)
        {
            t += step;
            p = (float)((double)((t - t_near) * pi_trange_inv) - M_PI);
            sample_at(v, org, dir, t, pos);
            bs = map_density(orc_volume_sample(v, pos), gp);
            ++ns;
            bound_lo = fminr(bound_lo, bs);
            bound_hi = fmaxr(bound_hi, bs);
            update_moments(m, trig, p, phase_step_size, fs, bs);
            fs = bs;
        }
        float last_step = 0.0f - p;
        sample_at(v, org, dir, t_far, pos);
        bs = map_density(orc_volume_sample(v, pos), gp);
        ++ns;
        bound_lo = fminr(bound_lo, bs);
        bound_hi = fmaxr(bound_hi, bs);
        update_moments(m, trig, 0.0f, last_step, fs, bs);

        if (gp->transform_to_bounds)
        {
            bound_lo = fmaxr(bound_lo - gp->transform_bound_eps, 0.0f);
            bound_hi = fminr(bound_hi + gp->transform_bound_eps, 1.0f);
            double inv_relaxed_bound = 1.0 / (double)(bound_hi - bound_lo);
            moments[0] = (float)((double)(trig[0] * 0.3183098861837907f - bound_lo) * inv_relaxed_bound);
            for (int j = 1; j < m; ++j)
                moments[j] = (float)((double)(trig[j] * 0.3183098861837907f) * inv_relaxed_bound);
        }
        else
        {
            for (int j = 0; j < m; ++j)
                moments[j] = trig[j] * 0.3183098861837907f;
        }
        if (n_samples)
            *n_samples = ns;
        return pack_bound(bound_lo, bound_hi);
    }
    if (n_samples)
        *n_samples = 0;
    if (gp->transform_to_bounds && gp->compact_image)
        return 0; /* moments deliberately left unwritten */
    for (int j = 0; j < m; ++j)
        moments[j] = 0.f;
    return 0;
}

/* generate_host's per-pixel loop (src/moment_raymarcher.cpp:121-145). The
 * optional compaction / prediction-coding stages are separate functions. */
void orc_generate_image(const orc_volume *v, const mray_camera *cam, int W, int H, int m, const mray_gen_params *p,
                        float *moments, uint32_t *bounds, int32_t *n_samples, int row0, int row_step, int nthreads)
{
    if (row_step < 1)
        row_step = 1;
    int nrows = (H - row0 + row_step - 1) / row_step;
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int r = 0; r < nrows; ++r)
    {
        int y = row0 + r * row_step;
        for (int x = 0; x < W; ++x)
        {
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            size_t pix = (size_t)y * W + x;
            int32_t ns = 0;
            uint32_t b = orc_generate_ray(v, org, dir, cam->znear, moments + pix * m, m, p, &ns);
            if (bounds)
                bounds[pix] = b;
            if (n_samples)
                n_samples[pix] = ns;
        }
    }
}

/* FP64 generate over the same discretisation: the reference's float t
 * sequence, node count and float sample values (trilinear in float, as
 * RegularGrid::sample), so that bounds are the same 16-bit values; the phases
 * of the reference's segments [p_k - dphi, p_k] are the reference's float
 * values, and the moment integral of each segment is evaluated in double with
 * exact sin/cos. */
uint32_t orc_generate_ray_f64(const orc_volume *v, const float org[3], const float dir[3], float znear,
                              double *moments, int m, const mray_gen_params *gp)
{
    double trig[512];
    for (int j = 0; j < m; ++j)
        trig[j] = 0.0;
    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, v->dmin, v->dmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    if (!(hit && (t_far > 0)))
    {
        if (!(gp->transform_to_bounds && gp->compact_image))
            for (int j = 0; j < m; ++j)
                moments[j] = 0.0;
        return 0;
    }
    float step = gp->step_size;
    float lo = 1.f, hi = 0.f;
    float t = t_near;
    float pi_trange_inv = (float)(M_PI / (double)(t_far - t_near));
    float p = (float)(-M_PI);
    float phase_step = (float)(M_PI * (double)(step / (t_far - t_near)));
    float pos[3];
    sample_at(v, org, dir, t, pos);
    float fs = map_density(orc_volume_sample(v, pos), gp);
    lo = fminr(lo, fs);
    hi = fmaxr(hi, fs);
    for (;;)
    {
        int last = !(t + step < t_far);
        float p_next, seg, bs;
        if (!last)
        {
            t += step;
            p_next = (float)((double)((t - t_near) * pi_trange_inv) - M_PI);
            seg = phase_step;
            sample_at(v, org, dir, t, pos);
        }
        else
        {
            p_next = 0.0f;
            seg = 0.0f - p;
            sample_at(v, org, dir, t_far, pos);
        }
        bs = map_density(orc_volume_sample(v, pos), gp);
        lo = fminr(lo, bs);
        hi = fmaxr(hi, bs);
        /* segment [pp, pn] = [p_next - seg, p_next] with the reference's float
         * end points; the line has the reference's slope (bs - fs) / seg and
         * passes through (pp, fs) (update_moments, src/raymarch_generate.h:37-40) */
        double pn = p_next, pp = (double)(p_next - seg), sg = pn - pp;
        double f0 = fs, f1 = bs;
        double g = (f1 - f0) / (double)seg;
        double f1l = f0 + g * sg; /* the line's value at pn */
        for (int j = 1; j < m; ++j)
        {
            double jd = j;
            /* Re[(g/j^2 + i(f0 + g(phi - pp))/j) e^{-i j phi}] from pp to pn */
            double cn = cos(jd * pn), sn = sin(jd * pn);
            double cp = cos(jd * pp), sp = sin(jd * pp);
            trig[j] += (g / (jd * jd)) * (cn - cp) + (f1l * sn - f0 * sp) / jd;
        }
        trig[0] += sg * (f0 + 0.5 * g * sg);
        fs = bs;
        p = p_next;
        if (last)
            break;
    }
    if (gp->transform_to_bounds)
    {
        lo = fmaxr(lo - gp->transform_bound_eps, 0.0f);
        hi = fminr(hi + gp->transform_bound_eps, 1.0f);
        double inv = 1.0 / ((double)hi - (double)lo);
        moments[0] = (trig[0] / M_PI - lo) * inv;
        for (int j = 1; j < m; ++j)
            moments[j] = trig[j] / M_PI * inv;
    }
    else
        for (int j = 0; j < m; ++j)
            moments[j] = trig[j] / M_PI;
    return pack_bound(lo, hi);
}

void orc_generate_image_f64(const orc_volume *v, const mray_camera *cam, int W, int H, int m,
                            const mray_gen_params *p, double *moments, uint32_t *bounds, int row0, int row_step,
                            int nthreads)
{
    if (row_step < 1)
        row_step = 1;
    int nrows = (H - row0 + row_step - 1) / row_step;
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int r = 0; r < nrows; ++r)
    {
        int y = row0 + r * row_step;
        for (int x = 0; x < W; ++x)
        {
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            size_t pix = (size_t)y * W + x;
            uint32_t b = orc_generate_ray_f64(v, org, dir, cam->znear, moments + pix * m, m, p);
            if (bounds)
                bounds[pix] = b;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* bounded MESE, FP32 (src/mese/MESE_dynamic.h, src/mese/complex_algebra.h) */
/* ------------------------------------------------------------------------ */

/* threads an OpenMP parallel region of this library gets by default */
int orc_omp_max_threads(void) { return omp_get_max_threads(); }

/* fast_atan (src/mese/complex_algebra.h:131-152) */
float orc_fast_atan(float tan)
{
    int negative = (tan < 0.0f);
    float abs_tan = negative ? (-tan) : tan;
    int greater_one = (abs_tan > 1.0f);
    float x = greater_one ? abs_tan : 1.0f;
    x = 1.0f / x;
    float y = greater_one ? 1.0f : abs_tan;
    x = x * y;
    y = x * x;
    float z = y * 0.020835f - 0.085133f;
    z = y * z + 0.180141f;
    z = y * z - 0.330299f;
    y = y * z + 0.999866f;
    z = y * x;
    z = z * -2.0f + 1.570796f;
    z = greater_one ? z : 0.0f;
    x = x * y + z;
    return negative ? (-x) : x;
}

/* trigonometricToExponentialMoments, real input (src/mese/MESE_dynamic.h:35-54) */
static void trig_to_exp_f(int m, cf *E, const float *t)
{
    float ph = 3.14159265f * t[0] - 1.57079633f;
    E[0] = cf_make(cosf(ph), sinf(ph));
    E[0] = cf_rmul(0.0795774715f, E[0]);
    for (int l = 1; l < m; ++l)
    {
        E[l] = cf_make(0.0f, 0.0f);
        for (int j = 0; j < l; ++j)
            E[l] = cf_add(E[l], cf_mul(cf_rmul((float)(l - j), E[j]), cf_make(t[l - j], 0.0f)));
        E[l] = cf_mul(E[l], cf_make(0.0f, (float)((double)(2.0f * M_PI) / (double)(float)l)));
    }
    E[0] = cf_rmul(2.0f, E[0]);
}

/* complex-input overload (src/mese/MESE_dynamic.h:56-75) */
static void trig_to_exp_fc(int m, cf *E, const cf *t)
{
    float ph = 3.14159265f * t[0].x - 1.57079633f;
    E[0] = cf_make(cosf(ph), sinf(ph));
    E[0] = cf_rmul(0.0795774715f, E[0]);
    for (int l = 1; l < m; ++l)
    {
        E[l] = cf_make(0.0f, 0.0f);
        for (int j = 0; j < l; ++j)
            E[l] = cf_add(E[l], cf_mul(cf_rmul((float)(l - j), E[j]), t[l - j]));
        E[l] = cf_mul(E[l], cf_make(0.0f, (float)((double)(2.0f * M_PI) / (double)(float)l)));
    }
    E[0] = cf_rmul(2.0f, E[0]);
}

/* exponentialToTrigonometricMoments (src/mese/MESE_dynamic.h:77-98) */
static void exp_to_trig_f(int m, cf *T, const cf *E)
{
    cf gamma_0 = cf_divr(E[0], 2.f);
    T[0] = cf_divr(gamma_0, 0.0795774715f);
    float ph = (float)atan2((double)T[0].y, (double)T[0].x);
    T[0] = cf_make((ph + 1.57079633f) / 3.14159265f, 0.0f);
    T[0].y = 0.0f;
    for (int l = 1; l < m; ++l)
    {
        T[l] = cf_make(0.0f, 0.0f);
        for (int j = 1; j < l; ++j)
        {
            cf c = cf_div(cf_make((float)(l - j), 0.0f), cf_rmul((float)l, gamma_0));
            T[l] = cf_add(T[l], cf_mul(cf_mul(c, E[j]), T[l - j]));
        }
        cf den = cf_mul(cf_make(0.0f, (float)(2.0f * M_PI)), gamma_0);
        T[l] = cf_sub(cf_div(E[l], den), T[l]);
    }
}

/* levinsonsAlgorithm (src/mese/MESE_dynamic.h:100-135) */
static void levinson_f(int m, cf *sol, const cf *col, cf *tmp)
{
    sol[0] = cf_make(1.0f / col[0].x, 0.0f);
    int invalid = 0;
    for (int j = 1; j < m; ++j)
    {
        cf u = cf_make(0.0f, 0.0f);
        if (!invalid)
            for (int k = 0; k < j; ++k)
                u = cf_add(u, cf_mul(sol[k], col[j - k]));
        float a2 = cf_abs2(u);
        if (a2 >= 1.0f)
        {
            u = cf_make(0.0f, 0.0f);
            a2 = 0;
            invalid = 1;
        }
        float factor = 1.0f / (1.0f - a2);
        for (int k = 1; k < j; ++k)
            tmp[k] = cf_conj(sol[j - k]);
        tmp[j] = cf_make(sol[0].x, 0.0f);
        sol[0] = cf_make(factor * sol[0].x, 0.0f);
        for (int k = 1; k < j; ++k)
            sol[k] = cf_rmul(factor, cf_sub(sol[k], cf_mul(tmp[k], u)));
        sol[j] = cf_rmul(factor, cf_rmul(-tmp[j].x, u));
    }
}

/* computeAutocorrelation (src/mese/MESE_dynamic.h:137-147) */
static void autocorr_f(int m, cf *out, const cf *s)
{
    for (int k = 0; k < m; ++k)
    {
        out[k] = cf_make(0.0f, 0.0f);
        for (int j = 0; j < m - k; ++j)
            out[k] = cf_add(out[k], cf_mul(s[j], cf_conj(s[j + k])));
    }
}

/* computeCorrelation (src/mese/MESE_dynamic.h:149-159) */
static void corr_f(int m, cf *out, const cf *lhs, const cf *rhs)
{
    for (int k = 0; k < m; ++k)
    {
        out[k] = cf_make(0.0f, 0.0f);
        for (int j = 0; j < m - k; ++j)
            out[k] = cf_add(out[k], cf_mul(lhs[k + j], rhs[j]));
    }
}

/* computeImaginaryCorrelation (src/mese/MESE_dynamic.h:161-171) */
static void imag_corr_f(int m, float *out, const cf *lhs, const cf *rhs)
{
    for (int k = 0; k < m; ++k)
    {
        out[k] = 0.0f;
        for (int j = 0; j < m - k; ++j)
            out[k] += lhs[k + j].x * rhs[j].y + lhs[k + j].y * rhs[j].x;
    }
}

/* prepareMESELagrange, real overload (src/mese/MESE_dynamic.h:194-211) */
static void prepare_mese_f(int m, float *out, const float *trig, cf *E, cf *P, cf *R)
{
    trig_to_exp_f(m, E, trig);
    levinson_f(m, P, E, R);
    for (int i = 0; i < m; ++i)
        P[i] = cf_rmul(6.28318531f, P[i]);
    autocorr_f(m, R, P);
    E[0] = cf_rmul(0.5f, E[0]);
    imag_corr_f(m, out, R, E);
    float nf = 1.0f / (3.14159265f * P[0].x);
    for (int i = 0; i < m; ++i)
        out[i] = nf * out[i];
}

/* prepareMESELagrange, complex overload (src/mese/MESE_dynamic.h:173-192) */
static void prepare_mese_fc(int m, cf *out, const cf *trig, cf *E, cf *P, cf *R)
{
    trig_to_exp_fc(m, E, trig);
    levinson_f(m, P, E, R);
    for (int i = 0; i < m; ++i)
        P[i] = cf_rmul(6.28318531f, P[i]);
    autocorr_f(m, R, P);
    E[0] = cf_rmul(0.5f, E[0]);
    corr_f(m, out, R, E);
    float nf = 1.0f / (3.14159265f * P[0].x);
    for (int i = 0; i < m; ++i)
        out[i] = cf_rmul(nf, cf_mul(cf_make(-0.0f, -1.0f), out[i]));
    out[0] = cf_make(out[0].x, 0.0f);
}

#define MAXM 512

void orc_trig_to_exp(int m, const float *trig, float *exp_c) { trig_to_exp_f(m, (cf *)exp_c, trig); }
void orc_trig_to_exp_c(int m, const float *trig_c, float *exp_c)
{
    trig_to_exp_fc(m, (cf *)exp_c, (const cf *)trig_c);
}
void orc_exp_to_trig(int m, const float *exp_c, float *trig_c) { exp_to_trig_f(m, (cf *)trig_c, (const cf *)exp_c); }
void orc_levinson(int m, const float *first_col_c, float *sol_c)
{
    cf tmp[MAXM];
    levinson_f(m, (cf *)sol_c, (const cf *)first_col_c, tmp);
}
void orc_autocorrelation(int m, const float *sig_c, float *out_c) { autocorr_f(m, (cf *)out_c, (const cf *)sig_c); }
void orc_correlation(int m, const float *lhs_c, const float *rhs_c, float *out_c)
{
    corr_f(m, (cf *)out_c, (const cf *)lhs_c, (const cf *)rhs_c);
}
void orc_imag_correlation(int m, const float *lhs_c, const float *rhs_c, float *out)
{
    imag_corr_f(m, out, (const cf *)lhs_c, (const cf *)rhs_c);
}
void orc_prepare_mese(int m, const float *trig, float *lagrange)
{
    cf E[MAXM], P[MAXM], R[MAXM];
    prepare_mese_f(m, lagrange, trig, E, P, R);
}
void orc_prepare_mese_c(int m, const float *trig_c, float *lagrange_c)
{
    cf E[MAXM], P[MAXM], R[MAXM];
    prepare_mese_fc(m, (cf *)lagrange_c, (const cf *)trig_c, E, P, R);
}

/* evaluateFourierSeries (src/mese/MESE_dynamic.h:213-226) */
float orc_evaluate_fourier(int m, float cx, float cy, const float *coeff)
{
    float result = 0;
    cf z = cf_make(cx, cy);
    cf pw = cf_make(1.f, 0.0f);
    for (int l = 1; l < m; ++l)
    {
        pw = cf_mul(pw, z);
        result += coeff[l] * pw.x;
    }
    return 2.0f * result + coeff[0];
}

/* evaluateMESELagrange (src/mese/MESE_dynamic.h:242-250) */
float orc_evaluate_mese(int m, float phase, const float *lagrange)
{
    float s = orc_evaluate_fourier(m, cosf(-phase), sinf(-phase), lagrange);
    return orc_fast_atan(s) * 0.318309886f + 0.5f;
}

/* prepare_moments_from_std_coding (src/moment_preparation.h:55-64) on a copy,
 * as MomentImageHost::prepare_moments does (src/moment_image.cpp:170-180). */
void orc_prepare_moments_std(int m, const float *moments, float bias, float *pmoments)
{
    float mc[MAXM];
    mc[0] = (1 - bias) * moments[0] + bias * 0.5f;
    for (int i = 1; i < m; ++i)
        mc[i] = (1 - bias) * moments[i] + bias * 0.0f;
    orc_prepare_mese(m, mc, pmoments);
}

/* ------------------------------------------------------------------------ */
/* bounded MESE, FP64 (same algorithm, double arithmetic, exact functions)   */
/* ------------------------------------------------------------------------ */
/* trigonometricToExponentialMoments (src/mese/MESE_dynamic.h:35-54) in
 * double: gamma_0 = e^{i(pi t_0 - pi/2)} / (4 pi), then the power series of
 * exp(2 pi i sum_{l>=1} t_l z^l), gamma_l = (2 pi i / l) sum_{j<l} (l - j)
 * t_{l-j} gamma_j, and gamma_0 doubled at the end. exp_c: m interleaved
 * complex doubles. */
void orc_trig_to_exp_f64(int m, const double *t, double *exp_c)
{
    cd *E = (cd *)exp_c;
    double ph = M_PI * t[0] - M_PI / 2;
    E[0] = cd_rmul(1.0 / (4.0 * M_PI), cd_make(cos(ph), sin(ph)));
    for (int l = 1; l < m; ++l)
    {
        cd acc = cd_make(0, 0);
        for (int j = 0; j < l; ++j)
            acc = cd_add(acc, cd_rmul((double)(l - j) * t[l - j], E[j]));
        E[l] = cd_mul(acc, cd_make(0.0, 2.0 * M_PI / l));
    }
    E[0] = cd_rmul(2.0, E[0]);
}

/* levinsonsAlgorithm (src/mese/MESE_dynamic.h:100-135) in double, with the
 * reference's guard: once |u|^2 >= 1 that and every later reflection
 * coefficient is 0. Solves T a = e_0 for the Hermitian Toeplitz matrix whose
 * first column is first_col (diagonal Re(first_col_0)). */
void orc_levinson_f64(int m, const double *first_col_c, double *sol_c)
{
    const cd *E = (const cd *)first_col_c;
    cd *P = (cd *)sol_c;
    cd tmp[MAXM];
    P[0] = cd_make(1.0 / E[0].x, 0.0);
    int invalid = 0;
    for (int j = 1; j < m; ++j)
    {
        cd u = cd_make(0, 0);
        if (!invalid)
            for (int k = 0; k < j; ++k)
                u = cd_add(u, cd_mul(P[k], E[j - k]));
        double a2 = cd_abs2(u);
        if (a2 >= 1.0)
        {
            u = cd_make(0, 0);
            a2 = 0;
            invalid = 1;
        }
        double f = 1.0 / (1.0 - a2);
        for (int k = 1; k < j; ++k)
            tmp[k] = cd_conj(P[j - k]);
        tmp[j] = cd_make(P[0].x, 0.0);
        P[0] = cd_make(f * P[0].x, 0.0);
        for (int k = 1; k < j; ++k)
            P[k] = cd_rmul(f, cd_sub(P[k], cd_mul(tmp[k], u)));
        P[j] = cd_rmul(f, cd_rmul(-tmp[j].x, u));
    }
}

/* computeAutocorrelation (src/mese/MESE_dynamic.h:137-147) in double:
 * out_k = sum_j s_j conj(s_{j+k}) */
void orc_autocorrelation_f64(int m, const double *sig_c, double *out_c)
{
    const cd *s = (const cd *)sig_c;
    cd *R = (cd *)out_c;
    for (int k = 0; k < m; ++k)
    {
        cd acc = cd_make(0, 0);
        for (int j = 0; j < m - k; ++j)
            acc = cd_add(acc, cd_mul(s[j], cd_conj(s[j + k])));
        R[k] = acc;
    }
}

/* computeImaginaryCorrelation (src/mese/MESE_dynamic.h:161-171) in double:
 * out_k = Im(sum_j lhs_{k+j} rhs_j) */
void orc_imag_correlation_f64(int m, const double *lhs_c, const double *rhs_c, double *out)
{
    const cd *L = (const cd *)lhs_c, *Rr = (const cd *)rhs_c;
    for (int k = 0; k < m; ++k)
    {
        double acc = 0.0;
        for (int j = 0; j < m - k; ++j)
            acc += L[k + j].x * Rr[j].y + L[k + j].y * Rr[j].x;
        out[k] = acc;
    }
}

/* prepareMESELagrange, real overload (src/mese/MESE_dynamic.h:194-211), in
 * double: exponential moments, Levinson, scale by 2 pi, autocorrelation,
 * gamma_0 halved, imaginary correlation, normalised by 1 / (pi P_0). */
void orc_prepare_mese_f64(int m, const double *t, double *out)
{
    cd E[MAXM], P[MAXM], R[MAXM];
    orc_trig_to_exp_f64(m, t, (double *)E);
    orc_levinson_f64(m, (const double *)E, (double *)P);
    for (int i = 0; i < m; ++i)
        P[i] = cd_rmul(2.0 * M_PI, P[i]);
    orc_autocorrelation_f64(m, (const double *)P, (double *)R);
    E[0] = cd_rmul(0.5, E[0]);
    orc_imag_correlation_f64(m, (const double *)R, (const double *)E, out);
    double nf = 1.0 / (M_PI * P[0].x);
    for (int k = 0; k < m; ++k)
        out[k] = nf * out[k];
}

/* prepare_moments_from_std_coding (src/moment_preparation.h:55-64) in
 * double: the bias lerp (m_0 towards 0.5, the others towards 0), then
 * prepareMESELagrange */
void orc_prepare_moments_std_f64(int m, const double *moments, double bias, double *out)
{
    double t[MAXM];
    t[0] = (1 - bias) * moments[0] + bias * 0.5;
    for (int i = 1; i < m; ++i)
        t[i] = (1 - bias) * moments[i];
    orc_prepare_mese_f64(m, t, out);
}

/* fast_atan (src/mese/complex_algebra.h:131-152) evaluated in double: the
 * polynomial is part of the reference's definition of the MESE density. */
static double fast_atan_f64(double tan)
{
    int negative = (tan < 0.0);
    double abs_tan = negative ? (-tan) : tan;
    int greater_one = (abs_tan > 1.0);
    double x = greater_one ? abs_tan : 1.0;
    x = 1.0 / x;
    double y = greater_one ? 1.0 : abs_tan;
    x = x * y;
    y = x * x;
    double z = y * (double)0.020835f - (double)0.085133f;
    z = y * z + (double)0.180141f;
    z = y * z - (double)0.330299f;
    y = y * z + (double)0.999866f;
    z = y * x;
    z = z * -2.0 + (double)1.570796f;
    z = greater_one ? z : 0.0;
    x = x * y + z;
    return negative ? (-x) : x;
}

double orc_evaluate_mese_f64(int m, double phase, const double *lagrange)
{
    double s = 0.0;
    for (int l = 1; l < m; ++l)
        s += lagrange[l] * cos(l * phase);
    s = 2.0 * s + lagrange[0];
    return fast_atan_f64(s) / M_PI + 0.5;
}

/* ------------------------------------------------------------------------ */
/* transfer function (src/transfer_function.cpp:46-119, .h:44-59)            */
/* ------------------------------------------------------------------------ */

/* read_tf + the value preintegrate keeps for every table row
 * (src/transfer_function.cpp:86-92,108-116). */
void orc_tf_from_rgba8(const uint8_t *rgba8, float step_size, float *tf)
{
    for (int i = 0; i < 256; ++i)
    {
        float r = powf(rgba8[i * 4 + 0] / 255.f, 2.2f);
        float g = powf(rgba8[i * 4 + 1] / 255.f, 2.2f);
        float b = powf(rgba8[i * 4 + 2] / 255.f, 2.2f);
        float a = rgba8[i * 4 + 3] / 255.f;
        a *= step_size;
        tf[i * 4 + 0] = r * a;
        tf[i * 4 + 1] = g * a;
        tf[i * 4 + 2] = b * a;
        tf[i * 4 + 3] = a;
    }
}

/* TransferFunction::sample on the row-constant table (src/transfer_function.h:44-59). */
void orc_tf_sample(const float *tf, float fs, float bs, float out[4])
{
    float ipx = fs * 256.0f, ipy = bs * 256.0f;
    float ax = ipx - floorf(ipx);
    float ay = ipy - floorf(ipy);
    int xl = (int)floorf(ipx);
    int x0 = iclampr(xl, 0, 255), x1 = iclampr(xl + 1, 0, 255);
    const float *ll = tf + x0 * 4, *lr = tf + x0 * 4, *rr = tf + x1 * 4, *rl = tf + x1 * 4;
    float w0 = (1.f - ax) * (1.f - ay), w1 = ax * (1.f - ay), w2 = (1.f - ax) * ay, w3 = ax * ay;
    for (int c = 0; c < 4; ++c)
        out[c] = w0 * ll[c] + w1 * lr[c] + w2 * rl[c] + w3 * rr[c];
}

static void tf_sample_f64(const double *tf, double fs, double bs, double out[4])
{
    double ipx = fs * 256.0, ipy = bs * 256.0;
    double ax = ipx - floor(ipx), ay = ipy - floor(ipy);
    int xl = (int)floor(ipx);
    int x0 = iclampr(xl, 0, 255), x1 = iclampr(xl + 1, 0, 255);
    (void)ax;
    for (int c = 0; c < 4; ++c)
        out[c] = (1.0 - ay) * tf[x0 * 4 + c] + ay * tf[x1 * 4 + c];
}

/* ------------------------------------------------------------------------ */
/* reconstruct (src/raymarch_reconstruct.h:127-200)                          */
/* ------------------------------------------------------------------------ */
static inline float mese_reconstruct(int m, const float *pm, float blo, float bhi, float p, int fourier)
{
    p = fclampr(p, (float)(-M_PI), 0.0f);
    float s;
    if (fourier)
    {
        /* TruncatedFourierReconstruction (src/reconstruction_techniques.h:43-64) */
        s = orc_evaluate_fourier(m, cosf(-p), sinf(-p), pm);
    }
    else
        s = orc_evaluate_mese(m, p, pm);
    return s * (bhi - blo) + blo;
}

void orc_reconstruct_ray(const float org[3], const float dir[3], const float bmin[3], const float bmax[3], float znear,
                         int m, const float *pm, float blo, float bhi, const float *tf, const mray_rec_params *rp,
                         float out[4], int32_t *n_steps)
{
    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, bmin, bmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    float step = rp->step_size;
    int fourier = rp->use_truncated_fourier;
    int ns = 0;
    if (hit && (t_far > 0))
    {
        float t = t_near;
        float p = (float)((double)((t - t_near) / (t_far - t_near)) * M_PI - M_PI);
        float phase_step = (float)(M_PI * (double)(step / (t_far - t_near)));
        float L[3] = {0.0f, 0.0f, 0.0f};
        float alpha = 0.0f, T = 1.0f, tau = 0.0f;
        float fs = mese_reconstruct(m, pm, blo, bhi, p, fourier);
        float bs;
        while (t <= t_far)
        {
            t += step;
            p += phase_step;
            ++ns;
            bs = mese_reconstruct(m, pm, blo, bhi, p, fourier);
            float rgba[4];
            orc_tf_sample(tf, fs, bs, rgba);
            float sigma_t = rgba[3];
            for (int c = 0; c < 3; ++c)
            {
                float Le = rp->medium_emission * rgba[c];
                L[c] += (Le + 0.0f * rp->medium_scattering) * T;
            }
            tau += sigma_t;
            T = (float)exp((double)(-tau));
            alpha = 1.0f - T;
            fs = bs;
            if (alpha >= rp->early_ray_termination)
            {
                alpha = 1.f;
                break;
            }
        }
        out[0] = L[0];
        out[1] = L[1];
        out[2] = L[2];
        out[3] = alpha;
    }
    else
        out[0] = out[1] = out[2] = out[3] = 0.f;
    if (n_steps)
        *n_steps = ns;
}

/* ------------------------------------------------------------------------ */
/* reference DVR (src/raymarch_reference.h:93-162)                           */
/* ------------------------------------------------------------------------ */
/* map_density (src/parameters.h:31-34) for ReferenceParameters */
static inline float map_density_ref(float s, const mray_ref_params *p)
{
    return fclampr((s - p->data_min) / (p->data_max - p->data_min), 0.0f, 1.0f);
}

/* reference_ray: march while t < t_far sampling the volume itself, the TF
 * lookup of (front, back) densities, the invalid-sample colour for samples
 * >= 1e10, emission-absorption compositing with T = exp(-tau) (double exp,
 * as the float argument promotes), early termination sets alpha = 1. */
void orc_reference_ray(const orc_volume *v, const float org[3], const float dir[3], float znear, const float *tf,
                       const mray_ref_params *p, float out[4], int32_t *n_steps)
{
    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, v->dmin, v->dmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    int ns = 0;
    if (hit && (t_far > 0))
    {
        float L[3] = {0.0f, 0.0f, 0.0f};
        float t = t_near, alpha = 0.0f, T = 1.0f, tau = 0.0f;
        float pos[3];
        sample_at(v, org, dir, t, pos);
        float fs = map_density_ref(orc_volume_sample(v, pos), p);
        while (t < t_far)
        {
            t += p->step_size;
            ++ns;
            sample_at(v, org, dir, t, pos);
            float s = orc_volume_sample(v, pos);
            float bs = map_density_ref(s, p);
            float rgba[4];
            orc_tf_sample(tf, fs, bs, rgba);
            if (s >= 1e10f)
            {
                rgba[0] = 0.0f;
                rgba[1] = 1.0f;
                rgba[2] = 0.0f;
                rgba[3] = 0.05f;
            }
            float sigma_t = rgba[3];
            for (int c = 0; c < 3; ++c)
            {
                float Le = p->medium_emission * rgba[c];
                L[c] += (Le + 0.0f * p->medium_scattering) * T;
            }
            tau += sigma_t;
            T = (float)exp((double)(-tau));
            alpha = 1.0f - T;
            fs = bs;
            if (alpha >= p->early_ray_termination)
            {
                alpha = 1.f;
                break;
            }
        }
        out[0] = L[0];
        out[1] = L[1];
        out[2] = L[2];
        out[3] = alpha;
    }
    else
        out[0] = out[1] = out[2] = out[3] = 0.f;
    if (n_steps)
        *n_steps = ns;
}

/* reference_host (src/moment_raymarcher.cpp:47-61): OpenMP over rows */
void orc_reference_image(const orc_volume *v, const mray_camera *cam, int W, int H, const float *tf,
                         const mray_ref_params *p, float *rgba, int32_t *n_steps, int row0, int row_step,
                         int nthreads)
{
    if (row_step < 1)
        row_step = 1;
    int nrows = (H - row0 + row_step - 1) / row_step;
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int r = 0; r < nrows; ++r)
    {
        int y = row0 + r * row_step;
        for (int x = 0; x < W; ++x)
        {
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            size_t pix = (size_t)y * W + x;
            orc_reference_ray(v, org, dir, cam->znear, tf, p, rgba + pix * 4, n_steps ? n_steps + pix : NULL);
        }
    }
}

void orc_reconstruct_image(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                           const float *data, const uint32_t *index, const uint32_t *bounds, int prediction_coded,
                           int coding_warp, const float *coding_params, const float *tf, const mray_rec_params *p,
                           float *rgba, int32_t *n_steps, int row0, int row_step, int nthreads)
{
    if (row_step < 1)
        row_step = 1;
    int nrows = (H - row0 + row_step - 1) / row_step;
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int r = 0; r < nrows; ++r)
    {
        int y = row0 + r * row_step;
        float pm[MAXM];
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            int nm = index ? (int)(index[pix + 1] - index[pix]) : m;
            size_t off = index ? index[pix] : pix * (size_t)m;
            float *out = rgba + pix * 4;
            if (nm == 0)
            {
                /* reconstruct_zero_ray (src/raymarch_reconstruct.h:127-131) */
                float c[4];
                orc_tf_sample(tf, 0.0f, 0.0f, c);
                out[0] = c[0];
                out[1] = c[1];
                out[2] = c[2];
                out[3] = 0.0f;
                if (n_steps)
                    n_steps[pix] = 0;
                continue;
            }
            if (p->use_truncated_fourier)
            {
                if (prediction_coded)
                {
                    /* get_trig_moments decode path is only needed for the Fourier
                     * variant of coded images; not on the MESE hot path. */
                    for (int l = 0; l < nm; ++l)
                        pm[l] = 0.0f;
                }
                else
                    for (int l = 0; l < nm; ++l)
                        pm[l] = data[off + l];
            }
            else if (prediction_coded)
                orc_prepare_moments_pred(nm, data + off, coding_warp, coding_params, pm);
            else
                orc_prepare_moments_std(nm, data + off, p->bias, pm);
            float blo = 0.f, bhi = 1.f;
            if (bounds)
            {
                blo = MRAY_BOUND_LO(bounds[pix]) / 65535.f;
                bhi = MRAY_BOUND_HI(bounds[pix]) / 65535.f;
            }
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            int32_t ns = 0;
            orc_reconstruct_ray(org, dir, bmin, bmax, cam->znear, nm, pm, blo, bhi, tf, p, out, &ns);
            if (n_steps)
                n_steps[pix] = ns;
        }
    }
}

/* FP64 reconstruct of one ray from given multipliers: the reference's t
 * loop and float phase sequence (it increments the phase in float),
 * densities, TF lerp and compositing in double with exact atan and exp. */
static void reconstruct_ray_f64(const float org[3], const float dir[3], const float bmin[3], const float bmax[3],
                                float znear, int m, const double *pm, double blo, double bhi, const double *tf,
                                const mray_rec_params *rp, double *out, int32_t *n_steps)
{
    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, bmin, bmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    int ns = 0;
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (hit && (t_far > 0))
    {
        float step = rp->step_size;
        float t = t_near;
        float p = (float)((double)((t - t_near) / (t_far - t_near)) * M_PI - M_PI);
        float phase_step = (float)(M_PI * (double)(step / (t_far - t_near)));
        double L[3] = {0, 0, 0}, tau = 0.0, T = 1.0, alpha = 0.0;
        double pc = dclampr(p, -M_PI, 0.0);
        double fs = orc_evaluate_mese_f64(m, pc, pm) * (bhi - blo) + blo;
        while (t <= t_far)
        {
            t += step;
            p += phase_step;
            ++ns;
            pc = dclampr(p, -M_PI, 0.0);
            double bs = orc_evaluate_mese_f64(m, pc, pm) * (bhi - blo) + blo;
            double c[4];
            tf_sample_f64(tf, fs, bs, c);
            for (int k = 0; k < 3; ++k)
                L[k] += rp->medium_emission * c[k] * T;
            tau += c[3];
            T = exp(-tau);
            alpha = 1.0 - T;
            fs = bs;
            if (alpha >= rp->early_ray_termination)
            {
                alpha = 1.0;
                break;
            }
        }
        out[0] = L[0];
        out[1] = L[1];
        out[2] = L[2];
        out[3] = alpha;
    }
    if (n_steps)
        *n_steps = ns;
}

void orc_reconstruct_image_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                               const double *moments, const uint32_t *bounds, const float *tf,
                               const mray_rec_params *rp, double *rgba, int32_t *n_steps, int nthreads)
{
    orc_reconstruct_rows_f64(cam, bmin, bmax, W, H, m, moments, bounds, tf, rp, rgba, n_steps, 0, 1, nthreads);
}

void orc_reconstruct_rows_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                              const double *moments, const uint32_t *bounds, const float *tf,
                              const mray_rec_params *rp, double *rgba, int32_t *n_steps, int row0, int row_step,
                              int nthreads)
{
    double tfd[1024];
    for (int i = 0; i < 1024; ++i)
        tfd[i] = tf[i];
    if (row_step < 1)
        row_step = 1;
    const int nrows = (H - row0 + row_step - 1) / row_step;
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int r = 0; r < nrows; ++r)
    {
        const int y = row0 + r * row_step;
        double pm[MAXM];
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            orc_prepare_moments_std_f64(m, moments + pix * (size_t)m, rp->bias, pm);
            double blo = bounds ? MRAY_BOUND_LO(bounds[pix]) / 65535.0 : 0.0;
            double bhi = bounds ? MRAY_BOUND_HI(bounds[pix]) / 65535.0 : 1.0;
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            reconstruct_ray_f64(org, dir, bmin, bmax, cam->znear, m, pm, blo, bhi, tfd, rp, rgba + pix * 4,
                                n_steps ? n_steps + pix : NULL);
        }
    }
}

/* d fast_atan_f64 / dt: derivative of the polynomial (not of atan). */
static double fast_atan_deriv_f64(double t)
{
    double at = t < 0.0 ? -t : t;
    int greater_one = at > 1.0;
    double x = greater_one ? 1.0 / at : at;
    double y = x * x;
    double P = (((double)0.020835f * y - (double)0.085133f) * y + (double)0.180141f) * y - (double)0.330299f;
    P = P * y + (double)0.999866f;
    double dP = ((4.0 * (double)0.020835f * y - 3.0 * (double)0.085133f) * y + 2.0 * (double)0.180141f) * y -
                (double)0.330299f;
    double dxp = P + 2.0 * y * dP;
    return greater_one ? dxp * y : dxp;
}

/* Reverse mode of reconstruct_ray_f64 on given multipliers: the march is run
 * forward storing every step, then swept backwards (the adjoint of
 * L += e c T, tau += sigma, T = exp(-tau), alpha = 1 - T), through the
 * transfer-function lookup (entries from fs, weight from frac(256 bs)) and
 * the density rho = fast_atan(S)/pi + 1/2 mapped to the bounds, to the
 * multipliers (dS/dlambda_l = c_l cos(l theta), c_0 = 1, c_l = 2) and to the
 * 256 x 4 table. Test infrastructure: the checker of the GPU adjoint. */
#define ORC_ADJ_MAXSTEPS 65536
void orc_reconstruct_adjoint_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H,
                                 int m, const float *lagrange, const uint32_t *bounds, const float *tf,
                                 const mray_rec_params *rp, const double *g_rgba, double *grad_lag, double *grad_tf,
                                 int nthreads)
{
    for (int i = 0; i < 1024; ++i)
        grad_tf[i] = 0.0;
#pragma omp parallel num_threads(set_threads(nthreads))
    {
        double gtf[1024];
        for (int i = 0; i < 1024; ++i)
            gtf[i] = 0.0;
        double *st = (double *)malloc(sizeof(double) * 8 * ORC_ADJ_MAXSTEPS);
        int *sx = (int *)malloc(sizeof(int) * 2 * ORC_ADJ_MAXSTEPS);
#pragma omp for
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
            {
                size_t pix = (size_t)y * W + x;
                double pm[MAXM];
                for (int l = 0; l < m; ++l)
                {
                    pm[l] = lagrange[pix * (size_t)m + l];
                    grad_lag[pix * (size_t)m + l] = 0.0;
                }
                double blo = bounds ? MRAY_BOUND_LO(bounds[pix]) / 65535.0 : 0.0;
                double bhi = bounds ? MRAY_BOUND_HI(bounds[pix]) / 65535.0 : 1.0;
                float org[3], dir[3];
                orc_pixel_ray(cam, W, H, x, y, org, dir);
                float t_near, t_far;
                int hit = orc_intersect_aabb(org, dir, bmin, bmax, &t_near, &t_far);
                t_near = fmaxr(t_near, cam->znear);
                if (!(hit && t_far > 0))
                    continue;
                /* forward, storing per step: theta, S, ay, T before, c[4] */
                float step = rp->step_size;
                float t = t_near;
                float p = (float)((double)((t - t_near) / (t_far - t_near)) * M_PI - M_PI);
                float phase_step = (float)(M_PI * (double)(step / (t_far - t_near)));
                double tau = 0.0, T = 1.0;
                double pc = dclampr(p, -M_PI, 0.0);
                double fs = orc_evaluate_mese_f64(m, pc, pm) * (bhi - blo) + blo;
                int K = 0, ert = 0;
                while (t <= t_far && K < ORC_ADJ_MAXSTEPS)
                {
                    t += step;
                    p += phase_step;
                    pc = dclampr(p, -M_PI, 0.0);
                    double S = 0.0;
                    for (int l = 1; l < m; ++l)
                        S += pm[l] * cos(l * pc);
                    S = 2.0 * S + pm[0];
                    double bs = (fast_atan_f64(S) / M_PI + 0.5) * (bhi - blo) + blo;
                    double ipx = fs * 256.0, ipy = bs * 256.0;
                    double ay = ipy - floor(ipy);
                    int xl = (int)floor(ipx);
                    int x0 = iclampr(xl, 0, 255), x1 = iclampr(xl + 1, 0, 255);
                    double *r = st + 8 * K;
                    r[0] = pc;
                    r[1] = S;
                    r[2] = ay;
                    r[3] = T;
                    for (int c = 0; c < 4; ++c)
                        r[4 + c] = (1.0 - ay) * tf[x0 * 4 + c] + ay * tf[x1 * 4 + c];
                    sx[2 * K] = x0;
                    sx[2 * K + 1] = x1;
                    ++K;
                    tau += r[7];
                    T = exp(-tau);
                    fs = bs;
                    if (1.0 - T >= rp->early_ray_termination)
                    {
                        ert = 1;
                        break;
                    }
                }
                /* reverse sweep */
                const double *g = g_rgba + pix * 4;
                double Tbar = ert ? 0.0 : -g[3]; /* alpha = 1 - T_final */
                double taubar = 0.0;
                double e = rp->medium_emission;
                for (int k = K - 1; k >= 0; --k)
                {
                    const double *r = st + 8 * k;
                    /* T_{k+1} = exp(-tau_{k+1}): the stored T of step k+1, or the final T */
                    double Tk1 = (k + 1 < K) ? st[8 * (k + 1) + 3] : T;
                    taubar += -Tk1 * Tbar;
                    double sigbar = taubar;
                    double cbar[4] = {g[0] * e * r[3], g[1] * e * r[3], g[2] * e * r[3], sigbar};
                    Tbar = g[0] * e * r[4] + g[1] * e * r[5] + g[2] * e * r[6];
                    int x0 = sx[2 * k], x1 = sx[2 * k + 1];
                    double rhobar = 0.0;
                    for (int c = 0; c < 4; ++c)
                    {
                        gtf[x0 * 4 + c] += cbar[c] * (1.0 - r[2]);
                        gtf[x1 * 4 + c] += cbar[c] * r[2];
                        rhobar += cbar[c] * 256.0 * ((double)tf[x1 * 4 + c] - (double)tf[x0 * 4 + c]);
                    }
                    double Sbar = rhobar * (bhi - blo) * fast_atan_deriv_f64(r[1]) / M_PI;
                    grad_lag[pix * (size_t)m] += Sbar;
                    for (int l = 1; l < m; ++l)
                        grad_lag[pix * (size_t)m + l] += 2.0 * Sbar * cos(l * r[0]);
                }
            }
#pragma omp critical
        for (int i = 0; i < 1024; ++i)
            grad_tf[i] += gtf[i];
        free(st);
        free(sx);
    }
}

/* FP64 evaluation of given (float) multipliers: isolates the reconstruct
 * march from the preparation. */
void orc_reconstruct_lagrange64_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H,
                                    int m, const double *lagrange, const uint32_t *bounds, const double *tf,
                                    const mray_rec_params *rp, double *rgba, int32_t *n_steps, int nthreads)
{
    orc_reconstruct_lagrange64_rows_f64(cam, bmin, bmax, W, H, m, lagrange, bounds, tf, rp, rgba, n_steps, 0, 1,
                                        nthreads);
}

void orc_reconstruct_lagrange64_rows_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W,
                                         int H, int m, const double *lagrange, const uint32_t *bounds,
                                         const double *tf, const mray_rec_params *rp, double *rgba, int32_t *n_steps,
                                         int row0, int row_step, int nthreads)
{
    if (row_step < 1)
        row_step = 1;
    const int nrows = (H - row0 + row_step - 1) / row_step;
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int r = 0; r < nrows; ++r)
    {
        const int y = row0 + r * row_step;
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            double blo = bounds ? MRAY_BOUND_LO(bounds[pix]) / 65535.0 : 0.0;
            double bhi = bounds ? MRAY_BOUND_HI(bounds[pix]) / 65535.0 : 1.0;
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            reconstruct_ray_f64(org, dir, bmin, bmax, cam->znear, m, lagrange + pix * (size_t)m, blo, bhi, tf, rp,
                                rgba + pix * 4, n_steps ? n_steps + pix : NULL);
        }
    }
}

void orc_reconstruct_lagrange_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H,
                                  int m, const float *lagrange, const uint32_t *bounds, const float *tf,
                                  const mray_rec_params *rp, double *rgba, int32_t *n_steps, int nthreads)
{
    double tfd[1024];
    for (int i = 0; i < 1024; ++i)
        tfd[i] = tf[i];
    double *lag = (double *)malloc(sizeof(double) * (size_t)W * H * m);
    for (size_t i = 0; i < (size_t)W * H * m; ++i)
        lag[i] = lagrange[i];
    orc_reconstruct_lagrange64_f64(cam, bmin, bmax, W, H, m, lag, bounds, tfd, rp, rgba, n_steps, nthreads);
    free(lag);
}

/* ------------------------------------------------------------------------ */
/* compaction (src/moment_compaction.h)                                      */
/* ------------------------------------------------------------------------ */

/* evaluate_unknown_moments (src/moment_compaction.h:35-60) */
static void eval_unknown_f(int num_moments, const cf *E, int max_moments, cf *newE, cf *lev)
{
    int mm = max_moments - 1;
    int n = num_moments - 1;
    levinson_f(num_moments, lev, E, newE);
    for (int i = 0; i < n + 1; ++i)
        lev[i] = cf_rmul(6.28318531f, lev[i]);
    for (int l = 0; l < n + 1; ++l)
        newE[l] = E[l];
    cf inv_lev0 = cf_div(cf_make(-1.0f, 0.0f), lev[0]);
    for (int k = 0; k < mm - n; ++k)
    {
        newE[n + k + 1] = cf_make(0.f, 0.0f);
        for (int j = 1; j < n + 1; ++j)
            newE[n + k + 1] = cf_add(newE[n + k + 1], cf_mul(newE[j + k], lev[n + 1 - j]));
        newE[n + k + 1] = cf_mul(newE[n + k + 1], inv_lev0);
    }
}

/* exponential_to_bounded_moments, including its use of exp_moments[l]
 * inside the j loop (src/moment_compaction.h:65-79). */
static void exp_to_bounded_f(int m, const cf *E, cf *B, int n)
{
    cf gamma_0 = cf_divr(E[0], 2.f);
    for (int l = n; l < m; ++l)
    {
        B[l] = cf_make(0.0f, 0.0f);
        for (int j = 0; j < l; ++j)
        {
            cf c = cf_div(cf_make((float)(l - j), 0.0f), cf_rmul((float)l, gamma_0));
            B[l] = cf_add(B[l], cf_mul(cf_mul(c, E[l]), B[l - j]));
        }
        cf den = cf_mul(cf_mul(cf_make(0.0f, 2.0f), cf_make((float)M_PI, 0.0f)), gamma_0);
        B[l] = cf_sub(cf_div(E[l], den), B[l]);
    }
}

/* rmse (src/moment_compaction.h:84-94) */
static float rmse_f(int n, int m, const cf *pred, const float *target)
{
    float sum = 0.0f;
    for (int k = n; k < m; ++k)
    {
        float diff = pred[k].x - target[k];
        sum += diff * diff;
    }
    return sqrtf(sum);
}

void orc_evaluate_unknown_moments(int n, const float *exp_c, int m, float *new_exp_c)
{
    cf lev[MAXM];
    eval_unknown_f(n, (const cf *)exp_c, m, (cf *)new_exp_c, lev);
}
void orc_exp_to_bounded_moments(int m, const float *exp_c, float *bounded_c, int n)
{
    exp_to_bounded_f(m, (const cf *)exp_c, (cf *)bounded_c, n);
}
float orc_rmse(int n, int m, const float *pred_c, const float *target) { return rmse_f(n, m, (const cf *)pred_c, target); }

/* evaluate_rrmse (src/moment_compaction.h:99-111) */
static float eval_rrmse_f(int n, int m, const float *moments, const cf *tE, cf *tExt, cf *tLev, float norm)
{
    eval_unknown_f(n, tE, m, tExt, tLev);
    cf *ext = tLev;
    for (int l = 0; l < n; ++l)
        ext[l] = cf_make(moments[l], 0.0f);
    exp_to_bounded_f(m, tExt, ext, n);
    return rmse_f(n, m, ext, moments) * norm;
}

/* determine_best_number_moments (src/moment_compaction.h:113-141) */
int orc_determine_best_number_moments(int m, const float *moments, float error_threshold)
{
    if (error_threshold == 0.0f)
        return m;
    cf tE[MAXM], tExt[MAXM], tLev[MAXM];
    float norm = 1.0f / moments[0];
    trig_to_exp_f(m, tE, moments);
    int n = m / 2;
    int s = m / 4;
    while (s > 0)
    {
        float rrmse = eval_rrmse_f(n, m, moments, tE, tExt, tLev, norm);
        float d = rrmse - error_threshold;
        if (((d < 0) ? -d : d) < 1e-8f)
            return n;
        else if (rrmse > error_threshold)
            n += s;
        else
            n -= s;
        s = s / 2;
    }
    return n;
}

/* compute_num_moments (src/moment_compaction.h:147-174) */
void orc_compute_num_moments_image(int W, int H, int m, const float *dense, const uint32_t *bounds,
                                   float error_threshold, uint32_t *counts, int nthreads)
{
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            const float *mo = dense + pix * (size_t)m;
            int is_zero;
            if (bounds)
            {
                float lo = MRAY_BOUND_LO(bounds[pix]) / 65535.f;
                float hi = MRAY_BOUND_HI(bounds[pix]) / 65535.f;
                is_zero = lo == hi;
            }
            else
                is_zero = mo[0] == 0.0f;
            int n = 0;
            if (!is_zero)
                n = orc_determine_best_number_moments(m, mo, error_threshold);
            counts[pix] = (uint32_t)n;
        }
}

/* MomentImageHost::compact (src/moment_image.cpp:69-94) */
uint32_t orc_compact(int W, int H, int m, const float *dense, const uint32_t *counts, uint32_t *index, float *out)
{
    uint32_t sum = 0;
    for (size_t pix = 0; pix < (size_t)W * H; ++pix)
    {
        uint32_t n = counts[pix];
        index[pix] = sum;
        if (out)
            for (uint32_t l = 0; l < n; ++l)
                out[sum + l] = dense[pix * (size_t)m + l];
        sum += n;
    }
    index[(size_t)W * H] = sum;
    return sum;
}

/* ------------------------------------------------------------------------ */
/* prediction coding (src/moment_prediction_coding.h)                        */
/* ------------------------------------------------------------------------ */

/* encode (src/moment_prediction_coding.h:123-159) */
static void encode_f(int m, const cf *E, cf *code, cf *P, cf *tmp)
{
    code[0] = E[0];
    P[0] = cf_make(1.0f / E[0].x, 0.0f);
    int invalid = 0;
    for (int l = 1; l < m; ++l)
    {
        cf u = cf_make(0.0f, 0.0f);
        for (int k = 0; k < l; ++k)
            u = cf_add(u, cf_mul(P[k], E[l - k]));
        code[l] = u;
        float a2 = cf_abs2(u);
        if (invalid || a2 >= 1.0f)
        {
            u = cf_make(0.0f, 0.0f);
            a2 = 0;
            invalid = 1;
        }
        float factor = 1.0f / (1.0f - a2);
        for (int k = 1; k < l; ++k)
            tmp[k] = cf_conj(P[l - k]);
        tmp[l] = cf_make(P[0].x, 0.0f);
        P[0] = cf_make(factor * P[0].x, 0.0f);
        for (int k = 1; k < l; ++k)
            P[k] = cf_rmul(factor, cf_sub(P[k], cf_mul(tmp[k], u)));
        P[l] = cf_rmul(factor, cf_rmul(-tmp[l].x, u));
    }
}

/* decode (src/moment_prediction_coding.h:164-206) */
static void decode_f(int m, const cf *code, cf *E, cf *P, cf *tmp)
{
    E[0] = code[0];
    P[0] = cf_make(1.0f / E[0].x, 0.0f);
    const float eps = 1e-6f;
    int invalid = 0;
    for (int l = 1; l < m; ++l)
    {
        cf u = code[l];
        float a2 = cf_abs2(u);
        if (invalid || a2 >= (1.0f - eps))
        {
            u = cf_make(0.0f, 0.0f);
            a2 = 0;
            invalid = 1;
        }
        cf g = cf_make(0.0f, 0.0f);
        for (int k = 1; k < l; ++k)
            g = cf_add(g, cf_mul(P[k], E[l - k]));
        E[l] = cf_mul(cf_div(cf_make((float)1.0, 0.0f), P[0]), cf_sub(u, g));
        float factor = 1.0f / (1.0f - a2);
        for (int k = 1; k < l; ++k)
            tmp[k] = cf_conj(P[l - k]);
        tmp[l] = cf_make(P[0].x, 0.0f);
        P[0] = cf_make(factor * P[0].x, 0.0f);
        for (int k = 1; k < l; ++k)
            P[k] = cf_rmul(factor, cf_sub(P[k], cf_mul(tmp[k], u)));
        P[l] = cf_rmul(factor, cf_rmul(-tmp[l].x, u));
    }
}

void orc_encode(int m, const float *exp_c, float *code_c)
{
    cf P[MAXM], tmp[MAXM];
    encode_f(m, (const cf *)exp_c, (cf *)code_c, P, tmp);
}
void orc_decode(int m, const float *code_c, float *exp_c, float *eval_poly_c)
{
    cf tmp[MAXM];
    decode_f(m, (const cf *)code_c, (cf *)exp_c, (cf *)eval_poly_c, tmp);
}

/* prediction_coding_warp / _inv_warp (src/moment_prediction_coding.h:47-67) */
static inline float pc_warp(int l, float x, int warp, const float *params)
{
    float w;
    if (warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED)
        w = (x + params[(l - 1) * 2]) / (params[(l - 1) * 2 + 1] * 2.f);
    else
        w = x * 0.5f + 0.5f;
    return fclampr(w, 0.0f, 1.0f);
}
static inline float pc_inv_warp(int l, float x, int warp, const float *params)
{
    if (warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED)
        return (x * params[(l - 1) * 2 + 1] * 2.f) - params[(l - 1) * 2];
    return x * 2.f - 1.f;
}

/* transform_quantization_real (src/moment_prediction_coding.h:81-93) */
void orc_transform_quantization_real(int m, const float *code_c, float c0, float *out)
{
    const cf *code = (const cf *)code_c;
    out[0] = c0;
    cf axis = cf_make(-code[0].y, code[0].x);
    float length_axis = sqrtf(cf_abs2(axis));
    cf inv_axis = cf_div(cf_make(length_axis, 0.0f), axis);
    for (int l = 1; l < m; ++l)
        out[l] = cf_mul(code[l], inv_axis).x;
}

/* transform_dequantization_real (src/moment_prediction_coding.h:98-118) */
void orc_transform_dequantization_real(int m, const float *in, float *code_c, int warp, const float *params)
{
    cf *out = (cf *)code_c;
    float c0 = in[0];
    float ph = 3.14159265f * c0 - 1.57079633f;
    out[0] = cf_make(cosf(ph), sinf(ph));
    out[0] = cf_rmul(2.0f * 0.0795774715f, out[0]);
    cf axis = cf_make(-out[0].y, out[0].x);
    float length_axis = sqrtf(cf_abs2(axis));
    cf factor = cf_divr(axis, length_axis);
    for (int l = 1; l < m; ++l)
    {
        float ql = pc_inv_warp(l, in[l], warp, params);
        out[l] = cf_rmul(ql, factor);
    }
}

/* find_coding_transform_parameters (src/coding_transform.h:37-68), including
 * the inclusion of pixels whose count equals l (these read the next pixel's
 * first value; the read past the end of the data for the very last pixel is
 * skipped here). */
void orc_find_coding_params(int W, int H, int m, const float *data, const uint32_t *index, float *params)
{
    size_t total = index[(size_t)W * H];
    for (int i = 0; i < (m - 1) * 2; ++i)
        params[i] = 0.0f;
    for (int l = 1; l < m; ++l)
    {
        float mn = 0.0f, mx = 0.0f;
        for (size_t pix = 0; pix < (size_t)W * H; ++pix)
        {
            int num = (int)(index[pix + 1] - index[pix]);
            if (num < l)
                continue;
            size_t at = index[pix] + (size_t)l;
            if (at >= total)
                continue;
            float value = fclampr(data[at], -1.0f, 1.0f);
            mn = (value < mn) ? value : mn; /* std::min(a, b) = (b < a) ? b : a */
            mx = (mx < value) ? value : mx; /* std::max(a, b) = (a < b) ? b : a */
        }
        params[(l - 1) * 2] = mn;
        params[(l - 1) * 2 + 1] = mx;
    }
    for (int pidx = 0; pidx < m - 1; ++pidx)
    {
        float d = params[pidx * 2 + 1] - params[pidx * 2];
        params[pidx * 2 + 1] = ((d < 0) ? -d : d) * 0.5f;
        float a = params[pidx * 2];
        params[pidx * 2] = (a < 0) ? -a : a;
    }
}

/* CodingParamType::quantize / dequantize (src/coding_params.h:56-78) */
void orc_quantize_coding_params(int m, const float *params, uint16_t *q)
{
    for (int i = 0; i < (m - 1) * 2; ++i)
        q[i] = (uint16_t)ceilf(fclampr(params[i], 0.0f, 1.0f) * 65535.f);
}
void orc_dequantize_coding_params(int m, const uint16_t *q, float *params)
{
    for (int i = 0; i < (m - 1) * 2; ++i)
        params[i] = q[i] / 65535.f;
}

/* MomentImageHost::prediction_encode (src/moment_image.cpp:96-152) */
void orc_prediction_encode_image(int W, int H, int m, float *data, const uint32_t *index, int warp, float bias,
                                 uint16_t *coding_params_q, int nthreads)
{
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int y = 0; y < H; ++y)
    {
        cf E[MAXM], code[MAXM], P[MAXM], tmp[MAXM];
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            size_t idx = index[pix];
            int nm = (int)(index[pix + 1] - index[pix]);
            if (nm <= 0)
                continue;
            float *d = data + idx;
            d[0] = (1 - bias) * d[0] + bias * 0.5f;
            for (int i = 1; i < nm; ++i)
                d[i] = (1 - bias) * d[i] + bias * 0.0f;
            trig_to_exp_f(nm, E, d);
            encode_f(nm, E, code, P, tmp);
            orc_transform_quantization_real(nm, (const float *)code, d[0], d);
        }
    }
    float params_f[2 * MAXM];
    for (int i = 0; i < 2 * (m - 1); ++i)
        params_f[i] = 0.0f;
    if (warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED)
    {
        float raw[2 * MAXM];
        orc_find_coding_params(W, H, m, data, index, raw);
        orc_quantize_coding_params(m, raw, coding_params_q);
        orc_dequantize_coding_params(m, coding_params_q, params_f);
    }
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            int nm = (int)(index[pix + 1] - index[pix]);
            if (nm <= 0)
                continue;
            float *d = data + index[pix];
            for (int l = 1; l < nm; ++l)
                d[l] = pc_warp(l, d[l], warp, params_f);
        }
}

/* prepare_moments_from_pred_coding (src/moment_preparation.h:32-53) */
void orc_prepare_moments_pred(int m, const float *moments, int warp, const float *params, float *pm)
{
    cf code[MAXM], E[MAXM], P[MAXM], tmp[MAXM];
    orc_transform_dequantization_real(m, moments, (float *)code, warp, params);
    decode_f(m, code, E, P, tmp);
    for (int i = 0; i < m; ++i)
        P[i] = cf_rmul(6.28318531f, P[i]);
    autocorr_f(m, tmp, P);
    E[0] = cf_rmul(0.5f, E[0]);
    imag_corr_f(m, pm, tmp, E);
    float nf = 1.0f / (3.14159265f * P[0].x);
    for (int i = 0; i < m; ++i)
        pm[i] = nf * pm[i];
}

/* ------------------------------------------------------------------------ */
/* SPH gather (src/volume.cpp:55-94, src/particle_grid.{cpp,cuh})            */
/* ------------------------------------------------------------------------ */

/* cubicSpline (src/cut/kernels.h:55-64) */
float orc_cubic_spline(float x)
{
    const float sigma = 0.5641895835477563f;
    if (x < 1.f)
        return sigma * (1.f - (3.f / 2.f) * x * x * (1.f - x * 0.5f));
    else if (x < 2.f)
        return sigma / 4.f * (2.f - x) * (2.f - x) * (2.f - x);
    return 0.0f;
}

typedef struct
{
    uint32_t sx, sy, sz; /* padded grid size */
    float minp[3];
    float inv_cell;
    float r2;
    uint32_t total;
} sph_grid;

static inline uint32_t next_multiple(uint32_t x, int mul) { return ((x - 1) / mul + 1) * mul; }

/* getIndexFromCell with 4^3 tiles (src/particle_grid.cuh:64-83) */
static inline uint32_t sph_cell_index(const sph_grid *g, int cx, int cy, int cz)
{
    int tsx = 4, tsy = 4, tsz = (g->sz == 1) ? 1 : 4;
    int widthInTiles = ((int)g->sx + tsx - 1) / tsx;
    int heightInTiles = ((int)g->sy + tsy - 1) / tsy;
    int tileX = cx / tsx, tileY = cy / tsy, tileZ = cz / tsz;
    int inX = cx % tsx, inY = cy % tsy, inZ = cz % tsz;
    return (uint32_t)((tileZ * heightInTiles * widthInTiles + tileY * widthInTiles + tileX) * (tsx * tsy * tsz) +
                      inZ * tsx * tsy + inY * tsx + inX);
}

static inline void sph_get_cell(const sph_grid *g, const float p[3], int c[3])
{
    for (int i = 0; i < 3; ++i)
        c[i] = (int)((p[i] - g->minp[i]) * g->inv_cell);
}

typedef struct
{
    uint32_t key, idx;
} kv;
static int kv_cmp(const void *a, const void *b)
{
    const kv *x = (const kv *)a, *y = (const kv *)b;
    if (x->key != y->key)
        return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

struct orc_sph
{
    sph_grid g;
    float *sx, *sy, *sz, *sv; /* particles in cell order */
    uint32_t *start, *count;  /* per cell: first particle (0xFFFFFFFF when empty), count */
    float inv_h;
};

/* Particles::create_grid (src/volume.cpp:55-81): the uniform grid of cell
 * size h over the particles' bounding box (createGridFromDomainCPU,
 * src/particle_grid.cpp:38-55; createGridFromDomain,
 * src/particle_grid.cuh:389-400), particles ordered by cell with a stable
 * order inside a cell (the reference's std::sort leaves that order
 * unspecified). */
orc_sph *orc_sph_create(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h)
{
    float mn[3] = {x[0], y[0], z[0]}, mx[3] = {x[0], y[0], z[0]};
    for (uint32_t i = 1; i < n; ++i)
    {
        const float q[3] = {x[i], y[i], z[i]};
        for (int k = 0; k < 3; ++k)
        {
            mn[k] = (q[k] < mn[k]) ? q[k] : mn[k];
            mx[k] = (mx[k] < q[k]) ? q[k] : mx[k];
        }
    }
    orc_sph *s = (orc_sph *)calloc(1, sizeof(orc_sph));
    sph_grid *g = &s->g;
    float cell = h;
    float inv_cellf = 1 / cell; /* Vec3 / scalar multiplies by the reciprocal (src/cut/cuda_math.h:467-472) */
    uint32_t gs[3];
    for (int k = 0; k < 3; ++k)
        gs[k] = (uint32_t)((mx[k] - mn[k]) * inv_cellf + 1);
    g->sx = next_multiple(gs[0], 4);
    g->sy = next_multiple(gs[1], 4);
    g->sz = (gs[2] > 1) ? next_multiple(gs[2], 4) : gs[2];
    for (int k = 0; k < 3; ++k)
        g->minp[k] = mn[k];
    g->inv_cell = (float)(1.0 / (double)cell);
    g->r2 = cell * cell;
    g->total = g->sx * g->sy * g->sz;

    kv *keys = (kv *)malloc(sizeof(kv) * n);
    for (uint32_t i = 0; i < n; ++i)
    {
        const float q[3] = {x[i], y[i], z[i]};
        int c[3];
        sph_get_cell(g, q, c);
        /* the largest coordinate can round to one cell past a padded size
         * (the reference then indexes past its cell array): clamp */
        const uint32_t s3[3] = {g->sx, g->sy, g->sz};
        for (int k = 0; k < 3; ++k)
            c[k] = c[k] < 0 ? 0 : (c[k] >= (int)s3[k] ? (int)s3[k] - 1 : c[k]);
        keys[i].key = sph_cell_index(g, c[0], c[1], c[2]);
        keys[i].idx = i;
    }
    qsort(keys, n, sizeof(kv), kv_cmp);
    s->sx = (float *)malloc(sizeof(float) * n * 4);
    s->sy = s->sx + n;
    s->sz = s->sy + n;
    s->sv = s->sz + n;
    for (uint32_t i = 0; i < n; ++i)
    {
        s->sx[i] = x[keys[i].idx];
        s->sy[i] = y[keys[i].idx];
        s->sz[i] = z[keys[i].idx];
        s->sv[i] = v[keys[i].idx];
    }
    s->start = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)g->total);
    s->count = (uint32_t *)calloc((size_t)g->total, sizeof(uint32_t));
    for (uint32_t c = 0; c < g->total; ++c)
        s->start[c] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; ++i)
    {
        uint32_t c = keys[i].key;
        if (i == 0 || c != keys[i - 1].key)
            s->start[c] = i;
        s->count[c]++;
    }
    free(keys);
    s->inv_h = 1.f / h;
    return s;
}

void orc_sph_destroy(orc_sph *s)
{
    if (!s)
        return;
    free(s->sx);
    free(s->start);
    free(s->count);
    free(s);
}

/* Particles::sample (src/volume.cpp:83-94): InterpolateParticlesOp
 * (src/particle_interpolate.h:35-58) over the particles of the 27 cells
 * around p (cells in z, y, x order, particles in cell order) closer than h,
 * normalised by the weight sum when it is not 0 */
float orc_sph_sample(const orc_sph *s, const float p[3])
{
    const sph_grid *g = &s->g;
    int c[3];
    sph_get_cell(g, p, c);
    float value = 0.0f, wsum = 0.0f;
    int z0 = (c[2] - 1 > 0) ? c[2] - 1 : 0, z1 = (c[2] + 2 < (int)g->sz) ? c[2] + 2 : (int)g->sz;
    int y0 = (c[1] - 1 > 0) ? c[1] - 1 : 0, y1 = (c[1] + 2 < (int)g->sy) ? c[1] + 2 : (int)g->sy;
    int x0 = (c[0] - 1 > 0) ? c[0] - 1 : 0, x1 = (c[0] + 2 < (int)g->sx) ? c[0] + 2 : (int)g->sx;
    for (int cz = z0; cz < z1; ++cz)
        for (int cy = y0; cy < y1; ++cy)
            for (int cx = x0; cx < x1; ++cx)
            {
                uint32_t ci = sph_cell_index(g, cx, cy, cz);
                if (s->start[ci] == 0xFFFFFFFFu)
                    continue;
                for (uint32_t q = s->start[ci]; q < s->start[ci] + s->count[ci]; ++q)
                {
                    float dx = p[0] - s->sx[q], dy = p[1] - s->sy[q], dz = p[2] - s->sz[q];
                    float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < g->r2)
                    {
                        /* InterpolateParticlesOp (src/particle_interpolate.h:45-51) */
                        float ex = s->sx[q] - p[0], ey = s->sy[q] - p[1], ez = s->sz[q] - p[2];
                        float w = orc_cubic_spline(sqrtf(ex * ex + ey * ey + ez * ez) * s->inv_h);
                        wsum += w;
                        value += w * s->sv[q];
                    }
                }
            }
    if (wsum != 0.0f)
        value /= wsum;
    return value;
}

/* Particles::sample evaluated at the voxel positions of an nx*ny*nz grid over
 * [dmin, dmax] (voxel i at dmin + i*(dmax-dmin)/n, the RegularGrid::sample
 * convention). */
void orc_particles_to_grid(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h,
                           const float dmin[3], const float dmax[3], int nx, int ny, int nz, float *grid,
                           int nthreads)
{
    orc_sph *s = orc_sph_create(x, y, z, v, n, h);
#pragma omp parallel for num_threads(set_threads(nthreads)) collapse(2)
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
            {
                float p[3];
                const int ii[3] = {ix, iy, iz};
                const int nn[3] = {nx, ny, nz};
                for (int k = 0; k < 3; ++k)
                    p[k] = dmin[k] + (float)ii[k] / (float)nn[k] * (dmax[k] - dmin[k]);
                grid[((size_t)iz * ny + iy) * nx + ix] = orc_sph_sample(s, p);
            }
    orc_sph_destroy(s);
}

/* ------------------------------------------------------------------------ */
/* quantization (src/moment_quantization.cpp)                               */
/* ------------------------------------------------------------------------ */

/* quantize (src/moment_quantization.cpp:33-50): clamp, normalise, scale by
 * 2^bits - 1 in float and truncate; 32 bits keep the clamped float's bits */
uint32_t orc_quantize_value(float value, int bits, float lo, float hi)
{
    value = fclampr(value, lo, hi);
    if (bits == 32)
    {
        uint32_t u;
        memcpy(&u, &value, sizeof u);
        return u;
    }
    const float norm = (value - lo) / (hi - lo);
    const uint32_t full = (uint32_t)((1u << bits) - 1u);
    const uint32_t q = (uint32_t)(norm * (float)(int32_t)full);
    return q & full;
}

/* dequantize (src/moment_quantization.cpp:52-63) */
float orc_dequantize_value(uint32_t q, int bits, float lo, float hi)
{
    if (bits == 32)
    {
        float f;
        memcpy(&f, &q, sizeof f);
        return fclampr(f, lo, hi);
    }
    const float norm = (float)q / (float)(int32_t)((1u << bits) - 1u);
    return norm * (hi - lo) + lo;
}

/* the value range of moment l: quantize_prediction_coding uses
 * PREDICTION_CODING_QUANTIZATION_MIN/MAX = [0, 1] (src/moment_prediction_coding.h:32-33)
 * for every moment, quantize [0, 1] for moment 0 and [-1/pi, 1/pi] (float
 * parameters from the double quotients) for the others
 * (src/moment_quantization.cpp:253-283, 300-329) */
static void quant_range(int l, int prediction_coded, float *lo, float *hi)
{
    if (l == 0 || prediction_coded)
    {
        *lo = 0.0f;
        *hi = 1.0f;
    }
    else
    {
        *lo = (float)(-1.0 / M_PI);
        *hi = (float)(1.0 / M_PI);
    }
}

/* quantize_prediction_coding / quantize of a compact image with pack_bits
 * (src/moment_quantization.cpp:65-111, 244-283, 300-329): pixels without
 * moments are skipped, the bits of each word are appended least significant
 * first, and a last partial byte is not written */
size_t orc_quantize_image(int W, int H, const float *data, const uint32_t *index, int prediction_coded,
                          const uint8_t *bits, uint8_t *out)
{
    size_t nbytes = 0;
    uint32_t buffer = 0;
    int cur = 0;
    for (size_t pix = 0; pix < (size_t)W * H; ++pix)
    {
        const uint32_t start = index[pix], n = index[pix + 1] - start;
        for (uint32_t l = 0; l < n; ++l)
        {
            float lo, hi;
            quant_range((int)l, prediction_coded, &lo, &hi);
            const uint32_t q = orc_quantize_value(data[start + l], bits[l], lo, hi);
            for (int b = 0; b < bits[l]; ++b)
            {
                if (q & (1u << b))
                    buffer |= 1u << cur;
                if (++cur == 8)
                {
                    out[nbytes++] = (uint8_t)buffer;
                    buffer = 0;
                    cur = 0;
                }
            }
        }
    }
    return nbytes;
}

/* dequantize_prediction_coding / dequantize with unpack_bits
 * (src/moment_quantization.cpp:113-163, 285-298, 331-360). Bits past the end
 * of the stream (the dropped partial byte) read as zero; the reference reads
 * past its word list there. */
void orc_dequantize_image(int W, int H, const uint32_t *index, int prediction_coded, const uint8_t *bits,
                          const uint8_t *in, size_t n_in, float *data)
{
    size_t bit = 0;
    for (size_t pix = 0; pix < (size_t)W * H; ++pix)
    {
        const uint32_t start = index[pix], n = index[pix + 1] - start;
        for (uint32_t l = 0; l < n; ++l)
        {
            uint32_t q = 0;
            for (int b = 0; b < bits[l]; ++b, ++bit)
                if ((bit >> 3) < n_in && (in[bit >> 3] >> (bit & 7)) & 1u)
                    q |= 1u << b;
            float lo, hi;
            quant_range((int)l, prediction_coded, &lo, &hi);
            data[start + l] = orc_dequantize_value(q, bits[l], lo, hi);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* error bounds and uncertainty (src/raymarch_reconstruct.h:203-367,        */
/* src/error_measures.h, src/transfer_function.cpp:99-248)                  */
/* ------------------------------------------------------------------------ */

/* TransferFunction::sample on a full 256 x 256 table (src/transfer_function.h:44-59):
 * the (x, y + 1) corner is weighted by a.x (1 - a.y) and (x + 1, y) by
 * (1 - a.x) a.y, as the reference writes it */
void orc_tf2d_sample(const float *tf, float fs, float bs, float out[4])
{
    float ipx = fs * 256.0f, ipy = bs * 256.0f;
    float ax = ipx - floorf(ipx), ay = ipy - floorf(ipy);
    int xl = (int)floorf(ipx), yl = (int)floorf(ipy);
    int x0 = iclampr(xl, 0, 255), x1 = iclampr(xl + 1, 0, 255);
    int y0 = iclampr(yl, 0, 255), y1 = iclampr(yl + 1, 0, 255);
    const float *ll = tf + ((size_t)y0 * 256 + x0) * 4, *lr = tf + ((size_t)y1 * 256 + x0) * 4;
    const float *rr = tf + ((size_t)y1 * 256 + x1) * 4, *rl = tf + ((size_t)y0 * 256 + x1) * 4;
    float w0 = (1.f - ax) * (1.f - ay), w1 = ax * (1.f - ay), w2 = (1.f - ax) * ay, w3 = ax * ay;
    for (int c = 0; c < 4; ++c)
        out[c] = w0 * ll[c] + w1 * lr[c] + w2 * rl[c] + w3 * rr[c];
}

/* read_tf (src/transfer_function.cpp:99-119) */
static void read_tf1d(const uint8_t *rgba8, float tf1d[256][4])
{
    for (int i = 0; i < 256; ++i)
    {
        tf1d[i][0] = powf(rgba8[i * 4 + 0] / 255.f, 2.2f);
        tf1d[i][1] = powf(rgba8[i * 4 + 1] / 255.f, 2.2f);
        tf1d[i][2] = powf(rgba8[i * 4 + 2] / 255.f, 2.2f);
        tf1d[i][3] = rgba8[i * 4 + 3] / 255.f;
    }
}

/* sample1D (src/transfer_function.cpp:32-44) */
static void sample1d(const float tf1d[256][4], float x, float out[4])
{
    float ip = x * 256.0f;
    float a = ip - floorf(ip);
    int xl = (int)floorf(ip);
    int i0 = iclampr(xl, 0, 255), i1 = iclampr(xl + 1, 0, 255);
    for (int c = 0; c < 4; ++c)
        out[c] = (1.f - a) * tf1d[i0][c] + a * tf1d[i1][c];
}

/* box_filter (src/transfer_function.cpp:152-185): entry (x, y) averages the
 * 1D function over the |y - x| + 2 points from x / 256 to y / 256 */
void orc_tf_box_filter(const uint8_t *rgba8, float step_size, float *table)
{
    float tf1d[256][4];
    read_tf1d(rgba8, tf1d);
#pragma omp parallel for
    for (int y = 0; y < 256; ++y)
        for (int x = 0; x < 256; ++x)
        {
            float sf = x / 256.0f, sb = y / 256.0f;
            int num_steps = abs(y - x) + 2;
            float dw = 1.0f / (float)num_steps;
            float rgba[4] = {0.f, 0.f, 0.f, 0.f};
            for (int i = 0; i < num_steps; ++i)
            {
                float w = (float)i / (float)(num_steps - 1);
                float sv = w * sb + (1 - w) * sf;
                float t[4];
                sample1d(tf1d, sv, t);
                float a = dw * t[3] * step_size;
                rgba[0] += t[0] * a;
                rgba[1] += t[1] * a;
                rgba[2] += t[2] * a;
                rgba[3] += a;
            }
            memcpy(table + ((size_t)y * 256 + x) * 4, rgba, sizeof rgba);
        }
}

/* normal_cdf (src/cut/gaussian.h:33-59): Abramowitz and Stegun 7.1.26 with
 * the reference's float / double promotions */
static float normal_cdf_std(float x)
{
    const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f, a5 = 1.061405429f,
                p = 0.3275911f;
    int sign = 1;
    if (x < 0)
        sign = -1;
    x = (float)((double)fabsf(x) / 1.41421356237);
    float t = (float)(1.0 / (1.0 + (double)(p * x)));
    float y = (float)(1.0 - (double)((((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * expf(-x * x)));
    return (float)(0.5 * (1.0 + (double)((float)sign * y)));
}

static float normal_cdf3(float x, float mean, float sigma) { return normal_cdf_std((x - mean) / sigma); }

/* gaussian_filter (src/transfer_function.cpp:187-231): entry (x, y) is the
 * 1D function weighted by the normal distribution of mean x / 256 and
 * standard deviation y / 256 over the 256 bins; row 0 is the function */
void orc_tf_gaussian_filter(const uint8_t *rgba8, float step_size, float *table)
{
    float tf1d[256][4];
    read_tf1d(rgba8, tf1d);
#pragma omp parallel for
    for (int y = 0; y < 256; ++y)
        for (int x = 0; x < 256; ++x)
        {
            float mean = x / 256.0f, sd = y / 256.0f;
            float rgba[4] = {0.f, 0.f, 0.f, 0.f};
            if (sd != 0.0f)
            {
                float delta = 1.0f / 256.0f;
                for (int i = 0; i < 256; ++i)
                {
                    float xi = i * delta;
                    float weight = normal_cdf3(xi + delta, mean, sd) - normal_cdf3(xi, mean, sd);
                    float a = weight * tf1d[i][3] * step_size;
                    rgba[0] += tf1d[i][0] * a;
                    rgba[1] += tf1d[i][1] * a;
                    rgba[2] += tf1d[i][2] * a;
                    rgba[3] += a;
                }
            }
            else
            {
                rgba[3] = tf1d[x][3] * step_size;
                rgba[0] = tf1d[x][0] * rgba[3];
                rgba[1] = tf1d[x][1] * rgba[3];
                rgba[2] = tf1d[x][2] * rgba[3];
            }
            memcpy(table + ((size_t)y * 256 + x) * 4, rgba, sizeof rgba);
        }
}

/* ErrorMax, ErrorRMSE and ErrorPercentile (the P^2 estimator of the
 * percentile, src/error_measures.h:29-231) */
typedef struct
{
    int type;
    float maxe, squares;
    int count;
    float q;
    int init;
    float h[5];
    int pos[5];
    float des[5];
} err_acc;

static void err_init(err_acc *e, int type, int percentile)
{
    memset(e, 0, sizeof *e);
    e->type = type;
    e->q = (float)percentile / 100.f;
}

static void sort5(float *v)
{
    /* sort_five_numbers (src/error_measures.h:58-98): ascending */
    for (int i = 1; i < 5; ++i)
        for (int j = i; j > 0 && v[j] < v[j - 1]; --j)
        {
            float t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
}

static float p2_parabolic(const err_acc *e, int i, int d)
{
    float t1 = d / (float)(e->pos[i + 1] - e->pos[i - 1]);
    float t2 = (e->pos[i] - e->pos[i - 1] + d) * (e->h[i + 1] - e->h[i]) / (float)(e->pos[i + 1] - e->pos[i]);
    float t3 = (e->pos[i + 1] - e->pos[i] - d) * (e->h[i] - e->h[i - 1]) / (float)(e->pos[i] - e->pos[i - 1]);
    return e->h[i] + t1 * (t2 + t3);
}

static float p2_linear(const err_acc *e, int i, int d)
{
    return e->h[i] + d * (e->h[i + d] - e->h[i]) / (float)(e->pos[i + d] - e->pos[i]);
}

static void err_update(err_acc *e, float ref, float pred)
{
    if (e->type == MRAY_ERROR_MAX)
    {
        e->maxe = fmaxr(e->maxe, fabsf(pred - ref));
        return;
    }
    if (e->type == MRAY_ERROR_RMSE)
    {
        e->count += 1;
        float d = pred - ref;
        e->squares += d * d;
        return;
    }
    float value = fabsf(pred - ref);
    if (e->init < 5)
    {
        e->h[e->init++] = value;
        if (e->init == 5)
        {
            sort5(e->h);
            for (int i = 0; i < 5; ++i)
                e->pos[i] = 1 + i;
            e->des[0] = 1.0f;
            e->des[1] = 1.0f + 2.0f * e->q;
            e->des[2] = 1.0f + 4.0f * e->q;
            e->des[3] = 3.0f + 2.0f * e->q;
            e->des[4] = 5.0f;
        }
        return;
    }
    int move = 0;
    if (value < e->h[0])
    {
        e->h[0] = value;
        move = 1;
    }
    else if (value < e->h[1])
        move = 1;
    else if (value < e->h[2])
        move = 2;
    else if (value < e->h[3])
        move = 3;
    else if (value < e->h[4])
        move = 4;
    else if (value > e->h[4])
    {
        e->h[4] = value;
        move = 4;
    }
    for (int k = move; k < 5; ++k)
        e->pos[k] += 1;
    e->des[1] += e->q / 2.0f;
    e->des[2] += e->q;
    e->des[3] += (1.0f + e->q) / 2.f;
    e->des[4] += 1.0f;
    /* adjust_heights: markers 1..3 (marker 4's desired position and position
     * both grow by exactly 1 per update, so it never moves) */
    for (int i = 1; i < 4; ++i)
    {
        float di = e->des[i] - e->pos[i];
        if ((di >= 1 && (e->pos[i + 1] - e->pos[i] > 1)) || (di <= -1 && (e->pos[i - 1] - e->pos[i] < -1)))
        {
            int sdi = (di > 0) ? 1 : -1;
            float qp = p2_parabolic(e, i, sdi);
            if ((e->h[i - 1] < qp) && (qp < e->h[i + 1]))
                e->h[i] = qp;
            else
                e->h[i] = p2_linear(e, i, sdi);
            e->pos[i] += sdi;
        }
    }
}

/* the estimate; fewer than 3 percentile samples read marker heights the
 * reference leaves uninitialised, here 0 */
static float err_value(const err_acc *e)
{
    if (e->type == MRAY_ERROR_MAX)
        return e->maxe;
    if (e->type == MRAY_ERROR_RMSE)
        return sqrtf(e->squares / (float)e->count);
    return e->h[2];
}

/* reconstruct_error (src/raymarch_reconstruct.h:323-367): from t_near in
 * steps of step_size while t <= t_far, the reconstruction at the phase of t
 * against the mapped volume sample at t */
float orc_error_bound_ray(const orc_volume *v, const float org[3], const float dir[3], float znear, int m,
                          const float *pm, float blo, float bhi, const mray_error_params *ep)
{
    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, v->dmin, v->dmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    if (!(hit && (t_far > 0)))
        return 0.0f;
    err_acc e;
    err_init(&e, ep->error_type, ep->error_percentile);
    mray_gen_params mp;
    memset(&mp, 0, sizeof mp);
    mp.data_min = ep->data_min;
    mp.data_max = ep->data_max;
    float t = t_near;
    while (t <= t_far)
    {
        float p = (float)((double)((t - t_near) / (t_far - t_near)) * M_PI - M_PI);
        float fs = mese_reconstruct(m, pm, blo, bhi, p, 0);
        float pos[3];
        sample_at(v, org, dir, t, pos);
        float ref = map_density(orc_volume_sample(v, pos), &mp);
        err_update(&e, ref, fs);
        t += ep->step_size;
    }
    return err_value(&e);
}

/* MomentImageHost::set_error_bound (src/moment_image.h:102-106) */
static uint16_t quantize_error_bound(float e) { return (uint16_t)(uint32_t)(int32_t)(e * 65535.f); }

/* reconstruct_error_host (src/moment_raymarcher.cpp:516-539): every pixel's
 * multipliers prepared (prediction-coded or standard), then the error of the
 * reconstruction along its ray; a pixel without moments reconstructs with
 * an empty series (0), where the reference reads the first entry of an
 * empty multiplier array */
void orc_error_bounds(const orc_volume *v, const mray_camera *cam, int W, int H, int m, const float *data,
                      const uint32_t *index, const uint32_t *bounds, int prediction_coded, int coding_warp,
                      const float *coding_params, const mray_error_params *ep, uint16_t *out, int nthreads)
{
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int y = 0; y < H; ++y)
    {
        float pm[MAXM];
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            int nm = index ? (int)(index[pix + 1] - index[pix]) : m;
            size_t off = index ? index[pix] : pix * (size_t)m;
            if (nm > 0 && prediction_coded)
                orc_prepare_moments_pred(nm, data + off, coding_warp, coding_params, pm);
            else if (nm > 0)
                orc_prepare_moments_std(nm, data + off, ep->bias, pm);
            else
                pm[0] = 0.0f;
            float blo = 0.f, bhi = 1.f;
            if (bounds)
            {
                blo = MRAY_BOUND_LO(bounds[pix]) / 65535.f;
                bhi = MRAY_BOUND_HI(bounds[pix]) / 65535.f;
            }
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            out[pix] = quantize_error_bound(orc_error_bound_ray(v, org, dir, cam->znear, nm > 0 ? nm : 1, pm, blo, bhi,
                                                                ep));
        }
    }
}

/* The density bounds at phase p (ErrorBoundMESEReconstruction and
 * MeanStdDevMESEReconstruction, src/reconstruction_techniques.h:99-135):
 * [clamp(rho - e), clamp(rho + e)] to the pixel's bounds, or (rho, e) */
static void density_bounds(int m, const float *pm, float blo, float bhi, float p, int fourier, float e,
                           int mean_std, float out[2])
{
    float rho = mese_reconstruct(m, pm, blo, bhi, p, fourier);
    if (mean_std)
    {
        out[0] = rho;
        out[1] = e;
    }
    else
    {
        out[0] = fclampr(rho - e, blo, bhi);
        out[1] = fclampr(rho + e, blo, bhi);
    }
}

/* reconstruct_ray with ErrorBoundInterpolator (src/reconstruction_techniques.h:140-156):
 * the preintegrated march over (1 - s) lower + s upper */
static void uncertainty_interp_ray(const float org[3], const float dir[3], const float bmin[3], const float bmax[3],
                                   float znear, int m, const float *pm, float blo, float bhi, float e, float s,
                                   const float *tf, const mray_rec_params *rp, float out[4])
{
    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, bmin, bmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    out[0] = out[1] = out[2] = out[3] = 0.f;
    if (!(hit && (t_far > 0)))
        return;
    float step = rp->step_size;
    int fourier = rp->use_truncated_fourier;
    float t = t_near;
    float p = (float)((double)((t - t_near) / (t_far - t_near)) * M_PI - M_PI);
    float phase_step = (float)(M_PI * (double)(step / (t_far - t_near)));
    float L[3] = {0.0f, 0.0f, 0.0f};
    float alpha = 0.0f, T = 1.0f, tau = 0.0f;
    float b[2];
    density_bounds(m, pm, blo, bhi, p, fourier, e, 0, b);
    float fs = (1 - s) * b[0] + s * b[1];
    while (t <= t_far)
    {
        t += step;
        p += phase_step;
        density_bounds(m, pm, blo, bhi, p, fourier, e, 0, b);
        float bs = (1 - s) * b[0] + s * b[1];
        float rgba[4];
        orc_tf_sample(tf, fs, bs, rgba);
        for (int c = 0; c < 3; ++c)
            L[c] += (rp->medium_emission * rgba[c] + 0.0f * rp->medium_scattering) * T;
        tau += rgba[3];
        T = (float)exp((double)(-tau));
        alpha = 1.0f - T;
        fs = bs;
        if (alpha >= rp->early_ray_termination)
        {
            alpha = 1.f;
            break;
        }
    }
    out[0] = L[0];
    out[1] = L[1];
    out[2] = L[2];
    out[3] = alpha;
}

/* reconstruct_ray_uncertainty_tf (src/raymarch_reconstruct.h:203-257): one
 * sample of the bounds per step, classified by the 2D uncertainty table */
static void uncertainty_conv_ray(const float org[3], const float dir[3], const float bmin[3], const float bmax[3],
                                 float znear, int m, const float *pm, float blo, float bhi, float e, int mean_std,
                                 const float *tf2d, const mray_rec_params *rp, float out[4])
{
    float t_near, t_far;
    int hit = orc_intersect_aabb(org, dir, bmin, bmax, &t_near, &t_far);
    t_near = fmaxr(t_near, znear);
    out[0] = out[1] = out[2] = out[3] = 0.f;
    if (!(hit && (t_far > 0)))
        return;
    float step = rp->step_size;
    int fourier = rp->use_truncated_fourier;
    float t = t_near;
    float p = (float)((double)((t - t_near) / (t_far - t_near)) * M_PI - M_PI);
    float phase_step = (float)(M_PI * (double)(step / (t_far - t_near)));
    float L[3] = {0.0f, 0.0f, 0.0f};
    float alpha = 0.0f, T = 1.0f, tau = 0.0f;
    while (t <= t_far)
    {
        t += step;
        p += phase_step;
        float b[2], rgba[4];
        density_bounds(m, pm, blo, bhi, p, fourier, e, mean_std, b);
        orc_tf2d_sample(tf2d, b[0], b[1], rgba);
        for (int c = 0; c < 3; ++c)
            L[c] += (rp->medium_emission * rgba[c] + 0.0f * rp->medium_scattering) * T;
        tau += rgba[3];
        T = (float)exp((double)(-tau));
        alpha = 1.0f - T;
        if (alpha >= rp->early_ray_termination)
        {
            alpha = 1.f;
            break;
        }
    }
    out[0] = L[0];
    out[1] = L[1];
    out[2] = L[2];
    out[3] = alpha;
}

/* reconstruct_uncertainty_interpolation_host / _convolution_host
 * (src/moment_raymarcher.cpp:563-712): mode 0 interpolates between the
 * bounds (s = bound_interpolation, tf the preintegrated table), mode 1
 * classifies (lower, upper) for MRAY_ERROR_MAX and (mean, error) otherwise
 * with the 2D uncertainty table. error_bounds: W*H uint16 (e / 65535). */
void orc_uncertainty_image(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                           const float *data, const uint32_t *index, const uint32_t *bounds,
                           const uint16_t *error_bounds, int prediction_coded, int coding_warp,
                           const float *coding_params, const float *tf, const mray_rec_params *p, int mode,
                           int error_type, float bound_interpolation, float *rgba, int nthreads)
{
#pragma omp parallel for num_threads(set_threads(nthreads))
    for (int y = 0; y < H; ++y)
    {
        float pm[MAXM];
        for (int x = 0; x < W; ++x)
        {
            size_t pix = (size_t)y * W + x;
            int nm = index ? (int)(index[pix + 1] - index[pix]) : m;
            size_t off = index ? index[pix] : pix * (size_t)m;
            float *out = rgba + pix * 4;
            if (nm == 0)
            {
                float c[4];
                if (mode == 0)
                    orc_tf_sample(tf, 0.0f, 0.0f, c);
                else
                    orc_tf2d_sample(tf, 0.0f, 0.0f, c);
                out[0] = c[0];
                out[1] = c[1];
                out[2] = c[2];
                out[3] = 0.0f;
                continue;
            }
            if (p->use_truncated_fourier)
                for (int l = 0; l < nm; ++l)
                    pm[l] = prediction_coded ? 0.0f : data[off + l];
            else if (prediction_coded)
                orc_prepare_moments_pred(nm, data + off, coding_warp, coding_params, pm);
            else
                orc_prepare_moments_std(nm, data + off, p->bias, pm);
            float blo = 0.f, bhi = 1.f;
            if (bounds)
            {
                blo = MRAY_BOUND_LO(bounds[pix]) / 65535.f;
                bhi = MRAY_BOUND_HI(bounds[pix]) / 65535.f;
            }
            const float e = (float)error_bounds[pix] / 65535.f;
            float org[3], dir[3];
            orc_pixel_ray(cam, W, H, x, y, org, dir);
            if (mode == 0)
                uncertainty_interp_ray(org, dir, bmin, bmax, cam->znear, nm, pm, blo, bhi, e, bound_interpolation, tf, p,
                                       out);
            else
                uncertainty_conv_ray(org, dir, bmin, bmax, cam->znear, nm, pm, blo, bhi, e,
                                     error_type != MRAY_ERROR_MAX, tf, p, out);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* entropy coding (src/moment_image_coding.cpp, src/entropy_coding/)        */
/* ------------------------------------------------------------------------ */

/* Amir Said's FastAC floating-point arithmetic coder (vendored by the
 * reference in src/entropy_coding/arithmetic_codec.{h,cpp}) with its adaptive
 * multi-symbol model, restated from the published algorithm:
 *   - the interval is [base, base + length) in doubles; a symbol s narrows it
 *     to [base + length d(s), base + length d(s + 1)), d(A) read as 1
 *     (the last symbol uses base + length); 2^-47 of leakage is taken off the
 *     length after every symbol;
 *   - base >= 1 propagates a carry into the bytes already written;
 *   - length <= 2^-16 shifts out two bytes (16 bits) of base at a time;
 *   - the model: counts start at 1, d(k) = sum_{i<k} count_i / total with
 *     total the sum of the counts, recomputed every `cycle` symbols; the
 *     counts are halved (rounding up) when total would pass 2^17, and cycle
 *     grows by 5/4 per update up to 8 (A + 6), starting from (A + 6) / 2.
 * Arithmetic_Codec::encode / decode (Adaptive_Data_Model):
 * src/entropy_coding/arithmetic_codec.cpp:445-519; renormalisation and carry
 * :76-112; start / stop :575-675; the model :838-919. */
#define AC_OUT 65536.0
#define AC_MINLEN (1.0 / 65536.0)
#define AC_LSB (1.0 / 16777216.0 / 16777216.0)
#define AC_LEAK (2.0 * AC_LSB)
#define AC_MAXCOUNT (1u << 17)

typedef struct
{
    unsigned n;      /* alphabet size */
    unsigned *count; /* symbol counts */
    double *dist;    /* cumulative distribution d(k) */
    unsigned total, cycle, until;
} ac_model;

static void ac_model_update(ac_model *M)
{
    M->total += M->cycle;
    if (M->total > AC_MAXCOUNT)
    {
        M->total = 0;
        for (unsigned k = 0; k < M->n; ++k)
        {
            M->count[k] = (M->count[k] + 1) >> 1;
            M->total += M->count[k];
        }
    }
    const double scale = 1.0 / (double)M->total;
    unsigned sum = 0;
    for (unsigned k = 0; k < M->n; ++k)
    {
        M->dist[k] = scale * (double)sum;
        sum += M->count[k];
    }
    M->cycle = (5 * M->cycle) >> 2;
    const unsigned cap = (M->n + 6) << 3;
    if (M->cycle > cap)
        M->cycle = cap;
    M->until = M->cycle;
}

static int ac_model_init(ac_model *M, unsigned n)
{
    M->n = n;
    M->count = (unsigned *)malloc(n * sizeof(unsigned));
    M->dist = (double *)malloc(n * sizeof(double));
    if (!M->count || !M->dist)
        return -1;
    for (unsigned k = 0; k < n; ++k)
        M->count[k] = 1;
    M->total = 0;
    M->cycle = n;
    ac_model_update(M);
    M->cycle = M->until = (n + 6) >> 1;
    return 0;
}

static void ac_model_free(ac_model *M)
{
    free(M->count);
    free(M->dist);
}

typedef struct
{
    double base, length;
    uint8_t *buf;
    size_t pos, cap;
    int overflow;
} ac_encoder;

static void ac_put(ac_encoder *E, unsigned b)
{
    if (E->pos < E->cap)
        E->buf[E->pos] = (uint8_t)b;
    else
        E->overflow = 1;
    ++E->pos;
}

static void ac_carry(ac_encoder *E)
{
    E->base -= 1.0;
    size_t p = E->pos - 1;
    while (p < E->cap && E->buf[p] == 0xFFu)
        E->buf[p--] = 0;
    if (p < E->cap)
        ++E->buf[p];
}

static void ac_encode(ac_encoder *E, ac_model *M, unsigned s)
{
    const double y = (s == M->n - 1) ? E->base + E->length : E->base + E->length * M->dist[s + 1];
    E->base += E->length * M->dist[s];
    E->length = y - E->base;
    if (E->base >= 1.0)
        ac_carry(E);
    if ((E->length -= AC_LEAK) <= AC_MINLEN)
    {
        do
        {
            E->base *= AC_OUT;
            const unsigned a = (unsigned)E->base;
            ac_put(E, a >> 8);
            ac_put(E, a & 0xFFu);
            E->base -= (double)a;
        } while ((E->length *= AC_OUT) <= AC_MINLEN);
    }
    ++M->count[s];
    if (--M->until == 0)
        ac_model_update(M);
}

static void ac_finish(ac_encoder *E)
{
    const unsigned a = (unsigned)(AC_OUT * E->base), b = (unsigned)(AC_OUT * (E->base + E->length));
    int last;
    if (b - a < 2)
    {
        E->base += 0.5 * AC_MINLEN;
        last = 3;
    }
    else if ((b >> 8) - (a >> 8) < 2)
    {
        E->base += AC_MINLEN;
        last = 2;
    }
    else
    {
        E->base += 256.0 * AC_MINLEN;
        last = 1;
    }
    if (E->base >= 1.0)
        ac_carry(E);
    do
    {
        E->base *= 256.0;
        const unsigned c = (unsigned)E->base;
        ac_put(E, c);
        E->base -= (double)c;
    } while (--last);
}

typedef struct
{
    double base, length, value;
    const uint8_t *buf;
    size_t pos, n;
} ac_decoder;

/* bytes past the end of the code read as zero */
static double ac_byte(const ac_decoder *D, size_t i) { return (double)(i < D->n ? D->buf[i] : 0u); }

static void ac_decoder_start(ac_decoder *D, const uint8_t *buf, size_t n)
{
    D->buf = buf;
    D->n = n;
    D->length = 1.0;
    D->base = D->value = 0.0;
    for (size_t k = 0; k < 6; ++k)
        D->value = 256.0 * D->value + AC_LSB * ac_byte(D, k);
    D->pos = 6;
}

static unsigned ac_decode(ac_decoder *D, ac_model *M)
{
    unsigned s = 0, n = M->n, mid = n >> 1;
    double x = D->base, y = D->base + D->length;
    const double sv = D->value + AC_LSB;
    do
    {
        const double z = D->base + D->length * M->dist[mid];
        if (z > sv)
        {
            y = z;
            n = mid;
        }
        else
        {
            x = z;
            s = mid;
        }
    } while ((mid = (s + n) >> 1) != s);
    D->base = x;
    D->length = y - x;
    if (D->base >= 1.0)
    {
        D->base -= 1.0;
        D->value -= 1.0;
    }
    if ((D->length -= AC_LEAK) <= AC_MINLEN)
    {
        do
        {
            D->base *= AC_OUT;
            const unsigned a = (unsigned)D->base;
            D->base -= (double)a;
            D->value = (AC_OUT * D->value - (double)a) + AC_LSB * (256.0 * ac_byte(D, D->pos) + ac_byte(D, D->pos + 1));
            D->pos += 2;
        } while ((D->length *= AC_OUT) <= AC_MINLEN);
    }
    ++M->count[s];
    if (--M->until == 0)
        ac_model_update(M);
    return s;
}

static void put_u32(uint8_t *p, uint32_t v) { memcpy(p, &v, 4); }
static uint32_t get_u32(const uint8_t *p)
{
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

/* entropy_encode (src/moment_image_coding.cpp:98-154, 202-217, write :73-96):
 * u32 count of non-empty pixels, their moment 0 quantized to bits[0] bits as
 * u16 words in pixel order, then for l = 1..m-1 a u32 byte count and the
 * arithmetic code of moment l of every pixel with more than l moments
 * (quantized to bits[l] bits, alphabet 2^bits[l]). Returns the section's
 * length; nothing is written past cap. -1 when a width is out of range. */
long orc_entropy_encode(int W, int H, int m, const float *data, const uint32_t *index, int prediction_coded,
                        const uint8_t *bits, uint8_t *out, size_t cap)
{
    const size_t npix = (size_t)W * H;
    for (int l = 1; l < m; ++l)
        if (bits[l] < 1 || bits[l] > 14)
            return -1;
    size_t pos = 4, nz = 0;
    for (size_t p = 0; p < npix; ++p)
        if (index[p + 1] > index[p])
        {
            const uint16_t q = (uint16_t)orc_quantize_value(data[index[p]], bits[0], 0.0f, 1.0f);
            if (pos + 2 <= cap)
                memcpy(out + pos, &q, 2);
            pos += 2;
            ++nz;
        }
    if (cap >= 4)
        put_u32(out, (uint32_t)nz);
    for (int l = 1; l < m; ++l)
    {
        float lo, hi;
        quant_range(l, prediction_coded, &lo, &hi);
        ac_model M;
        if (ac_model_init(&M, 1u << bits[l]))
            return -1;
        size_t nsym = 0;
        for (size_t p = 0; p < npix; ++p)
            nsym += (index[p + 1] - index[p]) > (uint32_t)l;
        ac_encoder E = {0.0, 1.0, NULL, 0, 4 * nsym + 16, 0};
        E.buf = (uint8_t *)malloc(E.cap);
        if (!E.buf)
            return -1;
        for (size_t p = 0; p < npix; ++p)
            if ((index[p + 1] - index[p]) > (uint32_t)l)
                ac_encode(&E, &M, orc_quantize_value(data[index[p] + l], bits[l], lo, hi));
        ac_finish(&E);
        if (pos + 4 <= cap)
            put_u32(out + pos, (uint32_t)E.pos);
        pos += 4;
        if (pos + E.pos <= cap)
            memcpy(out + pos, E.buf, E.pos);
        pos += E.pos;
        free(E.buf);
        ac_model_free(&M);
        if (E.overflow)
            return -1;
    }
    return (long)pos;
}

/* entropy_decode (src/moment_image_coding.cpp:156-200, 219-234, read :42-71)
 * of a section written by orc_entropy_encode into the compact data of the
 * image with the given index. Returns 0, or -1 when the section is shorter
 * than its headers say or a width is out of range. */
int orc_entropy_decode(int W, int H, int m, const uint8_t *in, size_t n_in, const uint32_t *index,
                       int prediction_coded, const uint8_t *bits, float *data)
{
    const size_t npix = (size_t)W * H;
    for (int l = 1; l < m; ++l)
        if (bits[l] < 1 || bits[l] > 14)
            return -1;
    if (n_in < 4)
        return -1;
    const size_t nz = get_u32(in);
    size_t pos = 4 + 2 * nz;
    if (pos > n_in)
        return -1;
    size_t k = 0;
    for (size_t p = 0; p < npix; ++p)
        if (index[p + 1] > index[p])
        {
            uint16_t q = 0;
            if (k < nz)
                memcpy(&q, in + 4 + 2 * k, 2);
            ++k;
            data[index[p]] = orc_dequantize_value(q, bits[0], 0.0f, 1.0f);
        }
    for (int l = 1; l < m; ++l)
    {
        if (pos + 4 > n_in)
            return -1;
        const size_t size = get_u32(in + pos);
        pos += 4;
        if (pos + size > n_in)
            return -1;
        float lo, hi;
        quant_range(l, prediction_coded, &lo, &hi);
        ac_model M;
        if (ac_model_init(&M, 1u << bits[l]))
            return -1;
        ac_decoder D;
        ac_decoder_start(&D, in + pos, size);
        for (size_t p = 0; p < npix; ++p)
            if ((index[p + 1] - index[p]) > (uint32_t)l)
                data[index[p] + l] = orc_dequantize_value(ac_decode(&D, &M), bits[l], lo, hi);
        ac_model_free(&M);
        pos += size;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* quantization curve (src/measure_coding_error.cpp)                        */
/* ------------------------------------------------------------------------ */

/* compute_reference / compute_coding_error with apply_quantization of one
 * moment (src/measure_coding_error.cpp:41-66, 92-182) */
void orc_coding_error(int W, int H, const float *data, const uint32_t *index, int warp, const float *params,
                      int q_index, int q_bits, float *ref, float *err)
{
    for (size_t pix = 0; pix < (size_t)W * H; ++pix)
    {
        const uint32_t start = index[pix];
        const int nm = (int)(index[pix + 1] - start);
        if (err)
            err[pix] = 0.0f;
        if (nm <= 0)
            continue;
        float in[MAXM];
        cf code[MAXM], E[MAXM], P[MAXM], tmp[MAXM], T[MAXM];
        memcpy(in, data + start, nm * sizeof(float));
        if (q_index >= 0 && nm > q_index)
        {
            const float p = (float)pow(2.0, (double)q_bits); /* std::pow(2.0f, b): exact */
            in[q_index] = floorf(in[q_index] * p) / p;
        }
        orc_transform_dequantization_real(nm, in, (float *)code, warp, params);
        decode_f(nm, code, E, P, tmp);
        exp_to_trig_f(nm, T, E);
        if (q_index < 0)
        {
            for (int l = 0; l < nm; ++l)
                ref[start + l] = T[l].x;
            continue;
        }
        float sum = 0.0f;
        for (int l = 0; l < nm; ++l)
        {
            const float d = ref[start + l] - T[l].x;
            sum += d * d;
        }
        err[pix] = sqrtf(sum) / ref[start];
    }
}

/* create_quantization_table with get_mean's float sum in pixel order
 * (src/measure_coding_error.cpp:184-192, 360-437) */
void orc_quantization_table(int W, int H, int m, const float *data, const uint32_t *index, int warp,
                            const float *params, int start_bits, uint8_t *table)
{
    const size_t npix = (size_t)W * H;
    float *ref = (float *)malloc((index[npix] + 1) * sizeof(float));
    float *err = (float *)malloc(npix * sizeof(float));
    size_t n = 0;
    for (size_t p = 0; p < npix; ++p)
        n += index[p + 1] > index[p];
    orc_coding_error(W, H, data, index, warp, params, -1, 0, ref, NULL);
#define MEAN_ERR(idx, b, out)                                                                                   \
    do                                                                                                          \
    {                                                                                                           \
        orc_coding_error(W, H, data, index, warp, params, (idx), (b), ref, err);                                \
        float mean_ = 0.0f;                                                                                     \
        for (size_t p = 0; p < npix; ++p)                                                                       \
            if (index[p + 1] > index[p])                                                                        \
                mean_ += err[p];                                                                                \
        out = mean_ / (float)n;                                                                                 \
    } while (0)
    for (int i = 0; i < m; ++i)
        table[i] = 0;
    table[0] = 16;
    if (m > 1)
        table[1] = (uint8_t)start_bits;
    float threshold = 0.0f;
    MEAN_ERR(1, start_bits, threshold);
    for (int idx = 2; idx < m; ++idx)
    {
        table[idx] = table[idx - 1];
        for (int b = table[idx - 1]; b >= 2; --b)
        {
            float e;
            MEAN_ERR(idx, b, e);
            if (e >= threshold)
                break;
            table[idx] = (uint8_t)b;
        }
    }
#undef MEAN_ERR
    free(ref);
    free(err);
}

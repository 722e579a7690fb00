// Thin extern "C" harness over the reference's own CPU code (TobiasRp/mray,
// src/...). TEST INFRASTRUCTURE ONLY. It is compiled together with the
// reference translation units where they lie under /root/reference (see
// oracle/Makefile) into oracle/_ref/libmrayref.so, and used in the build
// container to pin the C oracle and to write the golden vectors under
// tests/golden/. It contains no reference code: every function below builds
// the reference's own types and calls the reference's own functions.
#include "moment_raymarcher.h"
#include "cut/raytracing.h"
#include "raymarch_generate.h"
#include "raymarch_single_scattering_cache.h"
#include "raymarch_reconstruct.h"
#include "raymarch_reference.h"
#include "moment_compaction.h"
#include "moment_prediction_coding.h"
#include "moment_preparation.h"
#include "coding_transform.h"
#include "moment_quantization.h"
#include "measure_coding_error.h"
#include "moment_image_coding.h"
#include "moment_image_io.h"
#include "mese/MESE_4_10.h"

#include <cstdio>
#include <cstring>
#include <unistd.h>

#include "../include/mray_types.h"

namespace
{
cut::Camera make_cam(const mray_camera *c)
{
    cut::Camera cam;
    cam.orthographic = c->orthographic != 0;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            cam.transform[j][i] = c->transform[j * 4 + i];
    cam.position = Vec3f(c->position[0], c->position[1], c->position[2]);
    cam.right = Vec3f(c->right[0], c->right[1], c->right[2]);
    cam.up = Vec3f(c->up[0], c->up[1], c->up[2]);
    cam.dir = Vec3f(c->dir[0], c->dir[1], c->dir[2]);
    cam.scale = c->scale;
    cam.aspect = c->aspect;
    cam.znear = c->znear;
    return cam;
}

GenerationParameters make_gen(const mray_gen_params *p)
{
    GenerationParameters g;
    g.step_size = p->step_size;
    g.data_min = p->data_min;
    g.data_max = p->data_max;
    g.error_threshold = p->error_threshold;
    g.bias = p->bias;
    g.compact_image = p->compact_image != 0;
    g.prediction_coding = p->prediction_coding != 0;
    g.coding_warp = p->coding_warp;
    g.transform_to_bounds = p->transform_to_bounds != 0;
    g.transform_bound_eps = p->transform_bound_eps;
    return g;
}

ReconstructionParameters make_rec(const mray_rec_params *p)
{
    ReconstructionParameters r;
    r.bias = p->bias;
    r.medium_emission = p->medium_emission;
    r.medium_scattering = p->medium_scattering;
    r.early_ray_termination = p->early_ray_termination;
    r.data_min = p->data_min;
    r.data_max = p->data_max;
    r.step_size = p->step_size;
    r.use_truncated_fourier = p->use_truncated_fourier != 0;
    return r;
}

RegularGrid make_grid(const void *data, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax)
{
    Range<Vec3f> dom(Vec3f(dmin[0], dmin[1], dmin[2]), Vec3f(dmax[0], dmax[1], dmax[2]));
    size_t n = (size_t)nx * ny * nz;
    if (dtype == MRAY_DTYPE_U8)
    {
        vector<float> storage((n + 3) / 4, 0.0f);
        std::memcpy(storage.data(), data, n);
        return RegularGrid(cut::Vec3i(nx, ny, nz), dom, storage, RegularGrid::DataType::U8);
    }
    vector<float> storage((const float *)data, (const float *)data + n);
    return RegularGrid(cut::Vec3i(nx, ny, nz), dom, storage, RegularGrid::DataType::F32);
}

TransferFunction make_tf(const uint8_t *rgba8, float step)
{
    char path[] = "/tmp/mrayref_tfXXXXXX";
    int fd = mkstemp(path);
    FILE *f = fdopen(fd, "wb");
    fwrite(rgba8, 1, 1024, f);
    fclose(f);
    TransferFunction tf = read_from_disk(path, step);
    unlink(path);
    return tf;
}
} // namespace

namespace
{
// a compact image of the given counts and data, as moment_image_io reads one
MomentImageHost make_compact(int W, int H, const float *data, const uint32_t *index, int prediction_coded)
{
    MomentImageHost mi(W, H, 256, true, false);
    mi.index.assign(index, index + (size_t)W * H + 1);
    mi.data.assign(data, data + index[(size_t)W * H]);
    mi.is_compact = true;
    mi.prediction_code = prediction_coded != 0;
    return mi;
}
} // namespace

extern "C" {

// Full ::generate (src/moment_raymarcher.cpp:168-183 -> generate_host), which
// runs compute_num_moments / compact / prediction_encode as the parameters ask.
// Returns the number of floats in data; index (W*H+1) and coding params
// ((m-1)*2 uint16) are written when present.
long ref_generate(const void *vol, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                  const mray_camera *cam, int W, int H, int m, const mray_gen_params *p, float *data,
                  uint32_t *bounds, uint32_t *index, uint16_t *coding_params)
{
    auto grid = make_grid(vol, dtype, nx, ny, nz, dmin, dmax);
    auto gp = make_gen(p);
    MomentImageHost mi(W, H, m, gp.compact_image, gp.transform_to_bounds);
    mi.domain = grid.domain;
    generate(make_cam(cam), grid, mi, gp);
    std::memcpy(data, mi.data.data(), mi.data.size() * sizeof(float));
    if (bounds && mi.has_bounds())
        for (size_t i = 0; i < mi.bounds.size(); ++i)
            bounds[i] = (uint32_t)mi.bounds[i].lower | ((uint32_t)mi.bounds[i].upper << 16);
    if (index && !mi.index.empty())
        std::memcpy(index, mi.index.data(), mi.index.size() * sizeof(uint32_t));
    if (coding_params && mi.has_coding_params())
        for (size_t i = 0; i < mi.coding_params.size(); ++i)
        {
            coding_params[2 * i] = mi.coding_params[i].param1;
            coding_params[2 * i + 1] = mi.coding_params[i].param2;
        }
    return (long)mi.data.size();
}

// Per-pixel float RGBA of reconstruct_host's loop (src/moment_raymarcher.cpp:283-324)
// before the RGBA8 quantisation, plus the RGBA8 image of the real ::reconstruct.
void ref_reconstruct(const mray_camera *cam, const float *dmin, const float *dmax, int W, int H, int m,
                     const float *data, long n_data, const uint32_t *index, const uint32_t *bounds,
                     int prediction_coded, int coding_warp, const uint16_t *coding_params, const uint8_t *tf_rgba8,
                     float tf_step, const mray_rec_params *p, float *rgba, uint8_t *rgba8)
{
    MomentImageHost mi(W, H, m, index != nullptr, bounds != nullptr);
    mi.domain = Range<Vec3f>(Vec3f(dmin[0], dmin[1], dmin[2]), Vec3f(dmax[0], dmax[1], dmax[2]));
    mi.data.assign(data, data + n_data);
    if (index)
    {
        mi.index.assign(index, index + (size_t)W * H + 1);
        mi.is_compact = true;
    }
    if (bounds)
        for (size_t i = 0; i < (size_t)W * H; ++i)
        {
            mi.bounds[i].lower = (uint16_t)(bounds[i] & 0xFFFF);
            mi.bounds[i].upper = (uint16_t)(bounds[i] >> 16);
        }
    if (prediction_coded)
    {
        mi.prediction_code = true;
        mi.coding_warp = coding_warp;
        if (coding_warp == CODING_WARP_DEFAULT_TRANSFORMED)
            for (int l = 0; l < m - 1; ++l)
                mi.coding_params.push_back(CodingParamType{coding_params[2 * l], coding_params[2 * l + 1]});
    }
    auto tf = make_tf(tf_rgba8, tf_step);
    auto rp = make_rec(p);
    auto c = make_cam(cam);
    if (rgba)
    {
#pragma omp parallel for
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
            {
                Vec2f ndc(x / static_cast<float>(W) * 2.f - 1.f, y / static_cast<float>(H) * 2.f - 1.f);
                Vec4f col(0.0f);
                if (mi.get_num_moments(x, y) == 0)
                    col = reconstruct_zero_ray(tf);
                else if (rp.use_truncated_fourier)
                {
                    vector<float> pm(mi.get_num_moments(x, y));
                    mi.get_trig_moments(x, y, pm);
                    TruncatedFourierReconstructionDefault rec(pm.size(), pm.data(), mi.get_bounds(x, y));
                    col = reconstruct_ray(cut::get_eye_ray(c, ndc), mi.domain, c.znear, rec, tf, NoSingleScattering{},
                                          rp);
                }
                else
                {
                    auto pm = mi.prepare_moments(x, y, rp.bias);
                    MESEReconstructionDefault rec(pm.size(), pm.data(), mi.get_bounds(x, y));
                    col = reconstruct_ray(cut::get_eye_ray(c, ndc), mi.domain, c.znear, rec, tf, NoSingleScattering{},
                                          rp);
                }
                float *o = rgba + ((size_t)y * W + x) * 4;
                o[0] = col.x;
                o[1] = col.y;
                o[2] = col.z;
                o[3] = col.w;
            }
    }
    if (rgba8)
    {
        ColorImageHost img(W, H);
        reconstruct(c, mi, tf, img, rp);
        std::memcpy(rgba8, img.rgba.data(), (size_t)W * H * 4);
    }
}

// Reference direct volume rendering: per-pixel float RGBA of reference_host's
// loop (src/moment_raymarcher.cpp:47-61, reference_ray) before the RGBA8
// quantisation, plus the RGBA8 image of the real ::generate_reference
// (src/moment_raymarcher.cpp:74-81). Single scattering off.
void ref_reference_render(const void *vol, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                          const mray_camera *cam, int W, int H, const uint8_t *tf_rgba8, float tf_step,
                          const mray_ref_params *p, float *rgba, uint8_t *rgba8)
{
    auto grid = make_grid(vol, dtype, nx, ny, nz, dmin, dmax);
    auto tf = make_tf(tf_rgba8, tf_step);
    ReferenceParameters rp;
    rp.step_size = p->step_size;
    rp.medium_emission = p->medium_emission;
    rp.medium_scattering = p->medium_scattering;
    rp.early_ray_termination = p->early_ray_termination;
    rp.data_min = p->data_min;
    rp.data_max = p->data_max;
    rp.ss.enable = false;
    auto c = make_cam(cam);
    if (rgba)
    {
#pragma omp parallel for
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
            {
                Vec2f ndc(x / static_cast<float>(W) * 2.f - 1.f, y / static_cast<float>(H) * 2.f - 1.f);
                Vec4f col = reference_ray(grid, cut::get_eye_ray(c, ndc), c.znear, tf, SingleScatteringBruteForce{}, rp);
                float *o = rgba + ((size_t)y * W + x) * 4;
                o[0] = col.x;
                o[1] = col.y;
                o[2] = col.z;
                o[3] = col.w;
            }
    }
    if (rgba8)
    {
        ColorImageHost img(W, H);
        generate_reference(c, tf, grid, img, rp);
        std::memcpy(rgba8, img.rgba.data(), (size_t)W * H * 4);
    }
}

// Lagrange multipliers of MomentImageHost::prepare_moments (src/moment_image.cpp:154-183).
void ref_prepare_dense(int m, const float *moments, float bias, float *pm)
{
    MomentImageHost mi(1, 1, m, false, false);
    mi.data.assign(moments, moments + m);
    auto r = mi.prepare_moments(0, 0, bias);
    std::memcpy(pm, r.data(), m * sizeof(float));
}

// Single MESE building blocks (src/mese/MESE_dynamic.h) and the fixed-size
// versions the reference tests compare against (src/mese/MESE_4_10.h).
void ref_trig_to_exp(int m, const float *t, float *e) { trigonometricToExponentialMoments(m, (float_complex *)e, t); }
void ref_trig_to_exp_c(int m, const float *t, float *e)
{
    trigonometricToExponentialMoments(m, (float_complex *)e, (float_complex *)t);
}
void ref_exp_to_trig(int m, const float *e, float *t)
{
    exponentialToTrigonometricMoments(m, (float_complex *)t, (float_complex *)e);
}
void ref_levinson(int m, const float *col, float *sol)
{
    vector<float_complex> tmp(m);
    levinsonsAlgorithm(m, (float_complex *)sol, (float_complex *)col, tmp.data());
}
void ref_autocorrelation(int m, const float *s, float *o)
{
    computeAutocorrelation(m, (float_complex *)o, (float_complex *)s);
}
void ref_correlation(int m, const float *l, const float *r, float *o)
{
    computeCorrelation(m, (float_complex *)o, (float_complex *)l, (float_complex *)r);
}
void ref_imag_correlation(int m, const float *l, const float *r, float *o)
{
    computeImaginaryCorrelation(m, o, (float_complex *)l, (float_complex *)r);
}
void ref_prepare_mese(int m, const float *t, float *o)
{
    vector<float_complex> a(m), b(m), c(m);
    prepareMESELagrange(m, o, t, a.data(), b.data(), c.data());
}
void ref_prepare_mese_c(int m, const float *t, float *o)
{
    vector<float_complex> a(m), b(m), c(m);
    prepareMESELagrange(m, (float_complex *)o, (float_complex *)t, a.data(), b.data(), c.data());
}
float ref_evaluate_mese(int m, float phase, const float *pm) { return evaluateMESELagrange(m, phase, pm); }
float ref_fast_atan(float x) { return fast_atan(x); }

void ref_fixed_prepare4(const float *t, float *o) { prepareReflectanceSpectrumLagrange4(o, t); }
void ref_fixed_prepare10(const float *t, float *o) { prepareReflectanceSpectrumLagrange10(o, t); }
float ref_fixed_evaluate4(float p, const float *pm) { return evaluateReflectanceSpectrumLagrange4(p, pm); }
float ref_fixed_evaluate10(float p, const float *pm) { return evaluateReflectanceSpectrumLagrange10(p, pm); }
void ref_fixed_trig_to_exp4(const float *t, float *e) { trigonometricToExponentialMoments4((float_complex *)e, t); }
void ref_fixed_levinson4(const float *col, float *sol)
{
    levinsonsAlgorithm4((float_complex *)sol, (float_complex *)col);
}

// Compaction (src/moment_compaction.h)
int ref_best_num_moments(int m, const float *moments, float thr)
{
    vector<float_complex> a(m), b(m), c(m);
    return determine_best_number_moments(m, moments, thr, a.data(), b.data(), c.data());
}
void ref_eval_unknown(int n, const float *e, int m, float *ne)
{
    vector<float_complex> lev(m);
    evaluate_unknown_moments(n, (float_complex *)e, m, (float_complex *)ne, lev.data());
}
void ref_exp_to_bounded(int m, const float *e, float *b, int n)
{
    exponential_to_bounded_moments(m, (float_complex *)e, (float_complex *)b, n);
}

// Prediction coding (src/moment_prediction_coding.h)
void ref_encode(int m, const float *e, float *code)
{
    vector<float_complex> ep(m), tmp(m);
    encode(m, (float_complex *)e, (float_complex *)code, ep.data(), tmp.data());
}
void ref_decode(int m, const float *code, float *e, float *ep)
{
    vector<float_complex> tmp(m);
    decode(m, (float_complex *)code, (float_complex *)e, (float_complex *)ep, tmp.data());
}
void ref_transform_quant(int m, const float *code, float c0, float *out)
{
    transform_quantization_real(m, (float_complex *)code, c0, out);
}
void ref_transform_dequant(int m, const float *in, float *code, int warp, const float *params)
{
    transform_dequantization_real(m, in, (float_complex *)code, WarpParameters{warp, params});
}

// SPH: Particles::sample (src/volume.cpp:83-94) at the given points.
void ref_particles_sample(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h,
                          const float *dmin, const float *dmax, const float *pts, uint32_t npts, float *out)
{
    Range<Vec3f> dom(Vec3f(dmin[0], dmin[1], dmin[2]), Vec3f(dmax[0], dmax[1], dmax[2]));
    Particles parts(dom, vector<float>(x, x + n), vector<float>(y, y + n), vector<float>(z, z + n),
                    vector<float>(v, v + n), h);
#pragma omp parallel for
    for (long i = 0; i < (long)npts; ++i)
        out[i] = parts.sample(Vec3f(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}

// ::generate on a Particles (src/moment_raymarcher.cpp:168-183 ->
// generate_host over Particles::sample): dense moments and bounds.
void ref_generate_particles(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h,
                            const float *dmin, const float *dmax, const mray_camera *cam, int W, int H, int m,
                            const mray_gen_params *p, float *data, uint32_t *bounds)
{
    Range<Vec3f> dom(Vec3f(dmin[0], dmin[1], dmin[2]), Vec3f(dmax[0], dmax[1], dmax[2]));
    Particles parts(dom, vector<float>(x, x + n), vector<float>(y, y + n), vector<float>(z, z + n),
                    vector<float>(v, v + n), h);
    auto gp = make_gen(p);
    MomentImageHost mi(W, H, m, gp.compact_image, gp.transform_to_bounds);
    mi.domain = dom;
    generate(make_cam(cam), parts, mi, gp);
    std::memcpy(data, mi.data.data(), mi.data.size() * sizeof(float));
    if (bounds && mi.has_bounds())
        for (size_t i = 0; i < mi.bounds.size(); ++i)
            bounds[i] = (uint32_t)mi.bounds[i].lower | ((uint32_t)mi.bounds[i].upper << 16);
}

// moment_quantization::quantize_prediction_coding(mi, table) /
// moment_quantization::quantize(mi, bits) (src/moment_quantization.cpp:244-329).
// Returns the stream length in bytes.
long ref_quantize(int W, int H, const float *data, const uint32_t *index, int prediction_coded, const uint8_t *bits,
                  uint8_t *out)
{
    auto mi = make_compact(W, H, data, index, prediction_coded);
    vector<Byte> q;
    if (prediction_coded)
        q = moment_quantization::quantize_prediction_coding(mi, vector<Byte>(bits, bits + 256));
    else
        q = moment_quantization::quantize(mi, bits[0]);
    std::memcpy(out, q.data(), q.size());
    return (long)q.size();
}

// moment_quantization::dequantize_prediction_coding / dequantize
// (src/moment_quantization.cpp:285-360), on streams whose word count is whole
void ref_dequantize(int W, int H, const uint32_t *index, int prediction_coded, const uint8_t *bits,
                    const uint8_t *in, long n_in, float *data)
{
    vector<float> zeros(index[(size_t)W * H], 0.0f);
    auto mi = make_compact(W, H, zeros.data(), index, prediction_coded);
    vector<Byte> qs(in, in + n_in);
    if (prediction_coded)
        moment_quantization::dequantize_prediction_coding(mi, vector<Byte>(bits, bits + 256), qs);
    else
        moment_quantization::dequantize(mi, bits[0], qs);
    std::memcpy(data, mi.data.data(), mi.data.size() * sizeof(float));
}

// create_quantization_table (src/measure_coding_error.cpp:374-437) on a
// prediction-coded compact image; the table's first m entries
void ref_quantization_table(int W, int H, int m, const float *data, const uint32_t *index, int warp,
                            const uint16_t *coding_params, int start_bits, uint8_t *table)
{
    auto mi = make_compact(W, H, data, index, 1);
    mi.num_moments = m;
    mi.coding_warp = warp;
    mi.coding_params.resize(m - 1);
    for (int i = 0; i < m - 1; ++i)
    {
        mi.coding_params[i].param1 = coding_params[2 * i];
        mi.coding_params[i].param2 = coding_params[2 * i + 1];
    }
    create_quantization_table(mi, start_bits);
    auto t = moment_quantization::get_prediction_coding_quantization_table();
    std::memcpy(table, t.data(), m);
}

// entropy_encode (src/moment_image_coding.cpp:202-217) of a compact image of
// m moments with per-moment widths bits (m entries); returns the section's
// length, the bytes in out when it fits cap
long ref_entropy_encode(int W, int H, int m, const float *data, const uint32_t *index, int prediction_coded,
                        const uint8_t *bits, uint8_t *out, long cap)
{
    auto mi = make_compact(W, H, data, index, prediction_coded);
    mi.num_moments = m;
    auto bytes = entropy_encode(mi, vector<Byte>(bits, bits + m));
    if ((long)bytes.size() <= cap)
        std::memcpy(out, bytes.data(), bytes.size());
    return (long)bytes.size();
}

// entropy_decode (src/moment_image_coding.cpp:219-234) into compact data
void ref_entropy_decode(int W, int H, int m, const uint8_t *in, long n_in, const uint32_t *index,
                        int prediction_coded, const uint8_t *bits, float *data)
{
    vector<float> zeros(index[(size_t)W * H], 0.0f);
    auto mi = make_compact(W, H, zeros.data(), index, prediction_coded);
    mi.num_moments = m;
    // the decoder reads up to two bytes past each stream
    vector<Byte> bytes(in, in + n_in);
    entropy_decode(mi, vector<Byte>(bits, bits + m), bytes);
    std::memcpy(data, mi.data.data(), mi.data.size() * sizeof(float));
}

// write_moment_image (src/moment_image_io.cpp:515-530) of an image given by
// its arrays: compact with index (W*H+1) or dense (W*H*m floats); bounds
// (W*H packed lower | upper << 16), error_bounds (W*H u16) and
// coding_params (2(m-1) u16) when non-null. table (m entries) sets the
// prediction-coding quantization table first; the reference keeps it set
// for every later write in the process.
void ref_write_moment_image(const char *path, int W, int H, int m, int compact, int prediction_coded, int warp,
                            const float *dmin, const float *dmax, const uint32_t *index, const float *data,
                            long n_data, const uint32_t *bounds, const uint16_t *error_bounds,
                            const uint16_t *coding_params, const uint8_t *table, int compress, int bits,
                            int entropy)
{
    MomentImageHost mi(W, H, m, false, false);
    mi.domain = Range<Vec3f>(Vec3f(dmin[0], dmin[1], dmin[2]), Vec3f(dmax[0], dmax[1], dmax[2]));
    mi.coding_warp = warp;
    mi.prediction_code = prediction_coded != 0;
    mi.data.assign(data, data + n_data);
    if (compact)
    {
        mi.is_compact = true;
        mi.index.assign(index, index + (size_t)W * H + 1);
    }
    if (bounds)
    {
        mi.bounds.resize((size_t)W * H);
        for (size_t i = 0; i < mi.bounds.size(); ++i)
        {
            mi.bounds[i].lower = (uint16_t)(bounds[i] & 0xFFFFu);
            mi.bounds[i].upper = (uint16_t)(bounds[i] >> 16);
        }
    }
    if (error_bounds)
        mi.error_bounds.assign(error_bounds, error_bounds + (size_t)W * H);
    if (coding_params)
    {
        mi.coding_params.resize(m - 1);
        for (int i = 0; i < m - 1; ++i)
        {
            mi.coding_params[i].param1 = coding_params[2 * i];
            mi.coding_params[i].param2 = coding_params[2 * i + 1];
        }
    }
    if (table)
    {
        vector<Byte> t(255, 0);
        std::memcpy(t.data(), table, m);
        moment_quantization::set_prediction_coding_quantization_table(t);
    }
    write_moment_image(mi, path, compress != 0, (Byte)bits, entropy != 0);
}

// load_moment_image (src/moment_image_io.cpp:286-313): header fields
// (W, H, m, warp, compact, prediction coded, has bounds, has error bounds,
// has coding params, data length), then the arrays when the pointers are
// non-null (index W*H+1, data, bounds packed, coding params 2(m-1))
void ref_load_moment_image(const char *path, long *hdr, uint32_t *index, float *data, uint32_t *bounds,
                           uint16_t *coding_params)
{
    auto mi = load_moment_image(path);
    hdr[0] = mi.width;
    hdr[1] = mi.height;
    hdr[2] = mi.num_moments;
    hdr[3] = mi.coding_warp;
    hdr[4] = mi.is_compact;
    hdr[5] = mi.prediction_code;
    hdr[6] = mi.has_bounds();
    hdr[7] = mi.has_error_bounds();
    hdr[8] = mi.has_coding_params();
    hdr[9] = (long)mi.data.size();
    if (index && !mi.index.empty())
        std::memcpy(index, mi.index.data(), mi.index.size() * sizeof(uint32_t));
    if (data)
        std::memcpy(data, mi.data.data(), mi.data.size() * sizeof(float));
    if (bounds)
        for (size_t i = 0; i < mi.bounds.size(); ++i)
            bounds[i] = (uint32_t)mi.bounds[i].lower | ((uint32_t)mi.bounds[i].upper << 16);
    if (coding_params)
        for (size_t i = 0; i < mi.coding_params.size(); ++i)
        {
            coding_params[2 * i] = mi.coding_params[i].param1;
            coding_params[2 * i + 1] = mi.coding_params[i].param2;
        }
}

// read_from_disk_uncertainty (src/transfer_function.cpp:233-248) of a
// 256 x RGBA8 transfer function: the 256 x 256 x RGBA table
void ref_tf_uncertainty(const uint8_t *rgba8, float step, int error_type, float *table)
{
    char path[] = "/tmp/mrayref_utfXXXXXX";
    int fd = mkstemp(path);
    FILE *f = fdopen(fd, "wb");
    fwrite(rgba8, 1, 1024, f);
    fclose(f);
    TransferFunction tf = read_from_disk_uncertainty(path, step, error_type);
    unlink(path);
    std::memcpy(table, tf.data.data(), sizeof(float) * 4 * 256 * 256);
}

namespace
{
MomentImageHost make_image(const float *dmin, const float *dmax, int W, int H, int m, const float *data, long n_data,
                           const uint32_t *index, const uint32_t *bounds, const uint16_t *error_bounds,
                           int prediction_coded, int coding_warp, const uint16_t *coding_params)
{
    MomentImageHost mi(W, H, m, index != nullptr, bounds != nullptr);
    mi.domain = Range<Vec3f>(Vec3f(dmin[0], dmin[1], dmin[2]), Vec3f(dmax[0], dmax[1], dmax[2]));
    mi.data.assign(data, data + n_data);
    if (index)
    {
        mi.index.assign(index, index + (size_t)W * H + 1);
        mi.is_compact = true;
    }
    if (bounds)
        for (size_t i = 0; i < (size_t)W * H; ++i)
        {
            mi.bounds[i].lower = (uint16_t)(bounds[i] & 0xFFFF);
            mi.bounds[i].upper = (uint16_t)(bounds[i] >> 16);
        }
    if (error_bounds)
        mi.error_bounds.assign(error_bounds, error_bounds + (size_t)W * H);
    if (prediction_coded)
    {
        mi.prediction_code = true;
        mi.coding_warp = coding_warp;
        if (coding_warp == CODING_WARP_DEFAULT_TRANSFORMED)
            for (int l = 0; l < m - 1; ++l)
                mi.coding_params.push_back(CodingParamType{coding_params[2 * l], coding_params[2 * l + 1]});
    }
    return mi;
}
} // namespace

// reconstruct_error (src/moment_raymarcher.cpp:516-561, the host path) of a
// moment image against its volume: the per-pixel uint16 error bounds
void ref_error_bounds(const void *vol, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                      const mray_camera *cam, int W, int H, int m, const float *data, long n_data,
                      const uint32_t *index, const uint32_t *bounds, int prediction_coded, int coding_warp,
                      const uint16_t *coding_params, const mray_error_params *ep, uint16_t *out)
{
    auto grid = make_grid(vol, dtype, nx, ny, nz, dmin, dmax);
    auto mi = make_image(dmin, dmax, W, H, m, data, n_data, index, bounds, nullptr, prediction_coded, coding_warp,
                         coding_params);
    mi.add_error_bounds();
    ErrorReconstructionParameters p;
    p.error_type = ep->error_type;
    p.error_percentile = ep->error_percentile;
    p.bias = ep->bias;
    p.data_min = ep->data_min;
    p.data_max = ep->data_max;
    p.step_size = ep->step_size;
    reconstruct_error(make_cam(cam), grid, mi, p);
    std::memcpy(out, mi.error_bounds.data(), (size_t)W * H * sizeof(uint16_t));
}

// The uncertainty renderings (src/moment_raymarcher.cpp:563-712): float RGBA
// of the host loops before RGBA8 quantisation, and the RGBA8 image of the
// real functions. mode 0: reconstruct_uncertainty_interpolation with the
// preintegrated table of tf_rgba8 at tf_step; mode 1:
// reconstruct_uncertainty_convolution with read_from_disk_uncertainty's
// table at tf_step. Single scattering off.
void ref_uncertainty(const mray_camera *cam, const float *dmin, const float *dmax, int W, int H, int m,
                     const float *data, long n_data, const uint32_t *index, const uint32_t *bounds,
                     const uint16_t *error_bounds, int prediction_coded, int coding_warp,
                     const uint16_t *coding_params, const uint8_t *tf_rgba8, float tf_step, const mray_rec_params *p,
                     int mode, int error_type, float bound_interpolation, float *rgba, uint8_t *rgba8)
{
    auto mi = make_image(dmin, dmax, W, H, m, data, n_data, index, bounds, error_bounds, prediction_coded,
                         coding_warp, coding_params);
    TransferFunction tf;
    if (mode == 0)
        tf = make_tf(tf_rgba8, tf_step);
    else
        ref_tf_uncertainty(tf_rgba8, tf_step, error_type, reinterpret_cast<float *>(tf.data.data()));
    UncertaintyReconstructionParameters up;
    up.rec = make_rec(p);
    up.bound_interpolation = bound_interpolation;
    up.error_type = error_type;
    auto c = make_cam(cam);
    const auto &rp = up.rec;
    if (rgba)
    {
#pragma omp parallel for
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
            {
                Vec2f ndc(x / static_cast<float>(W) * 2.f - 1.f, y / static_cast<float>(H) * 2.f - 1.f);
                Vec4f col(0.0f);
                const float e = mi.get_error_bound(x, y);
                const auto ray = cut::get_eye_ray(c, ndc);
                if (mi.get_num_moments(x, y) == 0)
                    col = reconstruct_zero_ray(tf);
                else if (rp.use_truncated_fourier)
                {
                    vector<float> pm(mi.get_num_moments(x, y));
                    mi.get_trig_moments(x, y, pm);
                    TruncatedFourierReconstructionDefault rec(pm.size(), pm.data(), mi.get_bounds(x, y));
                    if (mode == 0)
                        col = reconstruct_ray(ray, mi.domain, c.znear,
                                              ErrorBoundInterpolator<ErrorBoundMESEReconstruction<
                                                  TruncatedFourierReconstructionDefault>>(
                                                  ErrorBoundMESEReconstruction<TruncatedFourierReconstructionDefault>(
                                                      rec, e),
                                                  bound_interpolation),
                                              tf, NoSingleScattering{}, rp);
                    else if (error_type == MAX_ERROR)
                        col = reconstruct_ray_uncertainty_tf(
                            ray, mi.domain, c.znear,
                            ErrorBoundMESEReconstruction<TruncatedFourierReconstructionDefault>(rec, e), tf,
                            NoSingleScattering{}, rp);
                    else
                        col = reconstruct_ray_uncertainty_tf(
                            ray, mi.domain, c.znear,
                            MeanStdDevMESEReconstruction<TruncatedFourierReconstructionDefault>(rec, e), tf,
                            NoSingleScattering{}, rp);
                }
                else
                {
                    auto pm = mi.prepare_moments(x, y, rp.bias);
                    MESEReconstructionDefault rec(pm.size(), pm.data(), mi.get_bounds(x, y));
                    if (mode == 0)
                        col = reconstruct_ray(
                            ray, mi.domain, c.znear,
                            ErrorBoundInterpolator<ErrorBoundMESEReconstruction<MESEReconstructionDefault>>(
                                ErrorBoundMESEReconstruction<MESEReconstructionDefault>(rec, e), bound_interpolation),
                            tf, NoSingleScattering{}, rp);
                    else if (error_type == MAX_ERROR)
                        col = reconstruct_ray_uncertainty_tf(
                            ray, mi.domain, c.znear, ErrorBoundMESEReconstruction<MESEReconstructionDefault>(rec, e),
                            tf, NoSingleScattering{}, rp);
                    else
                        col = reconstruct_ray_uncertainty_tf(
                            ray, mi.domain, c.znear, MeanStdDevMESEReconstruction<MESEReconstructionDefault>(rec, e),
                            tf, NoSingleScattering{}, rp);
                }
                float *o = rgba + ((size_t)y * W + x) * 4;
                o[0] = col.x;
                o[1] = col.y;
                o[2] = col.z;
                o[3] = col.w;
            }
    }
    if (rgba8)
    {
        ColorImageHost img(W, H);
        if (mode == 0)
            reconstruct_uncertainty_interpolation(c, mi, tf, img, up);
        else
            reconstruct_uncertainty_convolution(c, mi, tf, img, up);
        std::memcpy(rgba8, img.rgba.data(), (size_t)W * H * 4);
    }
}
}

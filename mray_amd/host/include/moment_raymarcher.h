// generate / reconstruct with the reference's signatures
// (src/moment_raymarcher.h:48-64), running on one MI355X through
// libmray_hip.so. Errors are thrown as std::runtime_error carrying the C
// ABI's message. The library holds no mutable global state: the kernel form
// (MRAY_FORM_MIXED, the default, or MRAY_FORM_PLAIN, the reference's FP32
// arithmetic) is the kernel_form field of the parameters of each call, so
// calls from different threads are independent.
#pragma once

#include "color_image.h"
#include "cut/camera.h"
#include "moment_image.h"
#include "parameters.h"
#include "transfer_function.h"
#include "volume.h"

// Ray-marches the volume into img (dense moments, bounds when img has
// bounds), then compacts and prediction-codes it as params ask.
void generate(cut::Camera cam, const RegularGrid &volume, MomentImageHost &img, const GenerationParameters &params);

// Particles are marched directly: every ray sample is Particles::sample (the
// cubic-spline weighted mean of the values of particles closer than the
// smoothing length, through the reference's cell grid).
void generate(cut::Camera cam, const Particles &volume, MomentImageHost &img, const GenerationParameters &params);

// Renders the moment image into RGBA8 (clamped and truncated, as the reference).
void reconstruct(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf, ColorImageHost &cimg,
                 const ReconstructionParameters &params);

// generate_reference (src/moment_raymarcher.h:40-41): the reference direct
// volume renderer into an RGBA8 image, and the float RGBA of a W x H image
void generate_reference(cut::Camera cam, const TransferFunction &tf, const RegularGrid &volume, ColorImageHost &img,
                        const ReferenceParameters &params);
void generate_reference(cut::Camera cam, const TransferFunction &tf, const RegularGrid &volume, int W, int H,
                        vector<Vec4f> &rgba, const ReferenceParameters &params);

// Float-RGBA overload (un-clamped), which the reference lacks.
void reconstruct(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf, vector<Vec4f> &rgba,
                 const ReconstructionParameters &params);

// reconstruct_error (src/moment_raymarcher.h:73-76): the error bound of every
// pixel's reconstruction against the volume (maximum, RMSE or a percentile
// of the difference along its ray) into mi.error_bounds (which it creates).
// Runs in the reference's FP32 arithmetic (MRAY_FORM_PLAIN multipliers):
// the bounds are those of the reference.
void reconstruct_error(cut::Camera cam, const RegularGrid &volume, MomentImageHost &mi,
                       const ErrorReconstructionParameters &params);
void reconstruct_error(cut::Camera cam, const Particles &volume, MomentImageHost &mi,
                       const ErrorReconstructionParameters &params);

// The uncertainty renderings (src/moment_raymarcher.h:78-86) of an image with
// error bounds, in the reference's FP32 arithmetic: interpolation between
// the bounds with the preintegrated tf (params.bound_interpolation), and
// convolution with an uncertainty tf of read_from_disk_uncertainty
// (params.error_type). Single scattering is not supported.
void reconstruct_uncertainty_interpolation(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf,
                                           ColorImageHost &cimg, const UncertaintyReconstructionParameters &params);
void reconstruct_uncertainty_convolution(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf,
                                         ColorImageHost &cimg, const UncertaintyReconstructionParameters &params);
// float-RGBA overloads (un-clamped), mode 0 interpolation, 1 convolution
void reconstruct_uncertainty(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf,
                             vector<Vec4f> &rgba, const UncertaintyReconstructionParameters &params, int mode);

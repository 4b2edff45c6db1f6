/*
 * libmvdecon_hip: MI355X (gfx950) implementation of the efficient-Bayesian
 * multiview deconvolution path of fiji/SPIM_Registration.
 *
 * Two groups of entry points, all plain C:
 *
 *  1. The JNA ABI the reference already binds.  The Java interfaces
 *     spim.process.cuda.CUDAStandardFunctions and CUDAFourierConvolution map
 *     these symbols by name (Java int -> int32, long -> int64, float[] ->
 *     float*, int[] -> int*, byte[] -> char*).  The library can be selected in
 *     the reference's native-library dialog unchanged.
 *
 *  2. A session API that keeps a whole deconvolution resident in HBM (views,
 *     weights, estimate, kernel spectra) and runs the MVDeconvolution
 *     iteration on the GPU.  It mirrors the reference operator API:
 *     MVDeconFFT (one view: image, weight, kernel, block size),
 *     MVDeconInput.init (kernel preparation per PSF type) and
 *     MVDeconvolution (initial estimate, iterations, Tikhonov, final mask).
 *
 * Arrays are x-fastest float32 volumes; host arrays are copied in and out,
 * the library never keeps or frees caller pointers.  Session functions return
 * 0 on success and a negative code on failure (message: mvd_last_error()).
 * A session belongs to one device; calls on one session must not overlap.
 */
#ifndef MVDECON_H
#define MVDECON_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. JNA ABI -------------------------------------------------------- */

/* Replaces CUDAStandardFunctions.getNumDevicesCUDA
 * (spim/process/cuda/CUDAStandardFunctions.java:41): number of GPUs, -1 if the
 * driver cannot be queried. */
int getNumDevicesCUDA(void);

/* Replaces CUDAStandardFunctions.getNameDeviceCUDA (CUDAStandardFunctions.java:42):
 * writes a NUL-terminated name into the caller's 256-byte buffer. */
void getNameDeviceCUDA(int devCUDA, char* name);

/* Replaces CUDAStandardFunctions.getMemDeviceCUDA (CUDAStandardFunctions.java:43):
 * total device memory in bytes, -1 on error. */
long long getMemDeviceCUDA(int devCUDA);

/* Replaces CUDAStandardFunctions.getFreeMemDeviceCUDA (CUDAStandardFunctions.java:44):
 * free device memory in bytes, -1 on error. */
long long getFreeMemDeviceCUDA(int devCUDA);

/* Replace CUDAStandardFunctions.getCUDAcomputeCapability{Major,Minor}Version
 * (CUDAStandardFunctions.java:36-37): the HIP device major/minor (9/5 on gfx950);
 * the reference only uses them to pick a default device. */
int getCUDAcomputeCapabilityMajorVersion(int devCUDA);
int getCUDAcomputeCapabilityMinorVersion(int devCUDA);

/* Replaces CUDAFourierConvolution.convolution3DfftCUDAInPlace
 * (spim/process/cuda/CUDAFourierConvolution.java:31), called per block at
 * deconv/MVDeconFFTThreads.java:86-89 and :109-112.
 * im: block of imDim = {nz, ny, nx} floats, x-fastest; kernel: {kz, ky, kx}.
 * Convolves im with kernel (true convolution, kernel centre at k/2 per axis,
 * circular over the block padded to an FFT-friendly size) and writes the
 * result into im.  Voxels at least k/2 from every block face -- the region
 * Block.pasteBlock copies back -- equal the linear convolution exactly; the
 * other voxels hold the circular result and are discarded by the caller.
 * On failure im is left unchanged and the error is printed to stderr (the
 * Java signature has no error channel). */
void convolution3DfftCUDAInPlace(float* im, int* imDim, float* kernel, int* kernelDim, int devCUDA);

/* Replaces CUDAFourierConvolution.convolution3DfftCUDA
 * (CUDAFourierConvolution.java:30; declared, never called by the reference).
 * Same result as the in-place call, returned in a new buffer the caller
 * releases with mvd_free(); NULL on failure. */
float* convolution3DfftCUDA(float* im, int* imDim, float* kernel, int* kernelDim, int devCUDA);
void mvd_free(void* p);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the block convolution of convolution3DfftCUDAInPlace on
 * device arrays, allocating nothing.  im: device block of imDim = {nz, ny,
 * nx} floats, overwritten with the result; kernel: device array of
 * kernelDim = {kz, ky, kx} floats; dims are host arrays, as in the JNA call.
 * Enqueued on `stream` (a hipStream_t, or NULL) and asynchronous: the result
 * is in im when the stream reaches it.  The kernel spectrum is recomputed on
 * every call.  Limit: the FFT length of each axis -- the smallest 5-smooth
 * size >= the block size, even along x -- at most 128 (blocks up to 128^3);
 * larger blocks are rejected with an error.  Scratch: one ~20 MiB device
 * array per GPU, held until the process ends, shared by every call on that
 * GPU: calls that overlap on one GPU race.  Returns 0, or -1 (invalid
 * arguments) / -2 (device error) with mvd_last_error() set. */
int mvd_convolution3d_inplace_noalloc(float* im, const int* imDim, const float* kernel, const int* kernelDim,
                                      int devCUDA, void* stream);

/* Replace CUDASeparableConvolution.convolve_{7,15,31,63,127}
 * (spim/process/cuda/CUDASeparableConvolution.java:34-38), the second native
 * hook (difference-of-Gaussian detection, via
 * CUDASeparableConvolutionFunctions.gauss, :123-217): in place, image of
 * imageW x imageH x imageD floats (x fastest), kernels of N floats centred at
 * N/2 (used where convolveX/Y/Z != 0, in that order); outofbounds 0 = zero,
 * 1 = outofboundsvalue, 2 = nearest border pixel.  Returns 1 on success, 0 on
 * failure (printed to stderr, image unchanged). */
int convolve_7(float* image, float* kernelX, float* kernelY, float* kernelZ, int imageW, int imageH, int imageD,
               int convolveX, int convolveY, int convolveZ, int outofbounds, float outofboundsvalue, int devCUDA);
int convolve_15(float* image, float* kernelX, float* kernelY, float* kernelZ, int imageW, int imageH, int imageD,
                int convolveX, int convolveY, int convolveZ, int outofbounds, float outofboundsvalue, int devCUDA);
int convolve_31(float* image, float* kernelX, float* kernelY, float* kernelZ, int imageW, int imageH, int imageD,
                int convolveX, int convolveY, int convolveZ, int outofbounds, float outofboundsvalue, int devCUDA);
int convolve_63(float* image, float* kernelX, float* kernelY, float* kernelZ, int imageW, int imageH, int imageD,
                int convolveX, int convolveY, int convolveZ, int outofbounds, float outofboundsvalue, int devCUDA);
int convolve_127(float* image, float* kernelX, float* kernelY, float* kernelZ, int imageW, int imageH, int imageD,
                 int convolveX, int convolveY, int convolveZ, int outofbounds, float outofboundsvalue, int devCUDA);
/* Replaces CUDASeparableConvolution.convolutionCPU (:42): kernels of
 * 2 * kernelR + 1 taps (NULL = axis skipped).  Runs the same passes on GPU 0
 * (this library has no CPU path); errors are printed. */
void convolutionCPU(float* image, float* kernelX, float* kernelY, float* kernelZ, int kernelRX, int kernelRY,
                    int kernelRZ, int imageW, int imageH, int imageD, int outofbounds, float outofboundsvalue);

/* ---- 2. Session API ----------------------------------------------------- */

typedef struct mvd_session mvd_session;

/* MVDeconFFT.PSFTYPE (deconv/MVDeconFFT.java:49), same order. */
enum {
  MVD_PSF_OPTIMIZATION_II = 0,
  MVD_PSF_OPTIMIZATION_I = 1,
  MVD_PSF_EFFICIENT_BAYESIAN = 2,
  MVD_PSF_INDEPENDENT = 3
};

/* How the weights given to mvd_add_view are used
 * (ProcessForDeconvolution.WeightType, deconv/ProcessForDeconvolution.java:221-243). */
enum {
  MVD_WEIGHTS_NONE = 0,        /* w = 1 everywhere (NO_WEIGHTS) */
  MVD_WEIGHTS_AS_GIVEN = 1,    /* weights used as passed (already normalised) */
  MVD_WEIGHTS_PRECOMPUTED = 2, /* normalised in place: w_v /= sum_u w_u (WeightNormalizer.ApplyDirectly) */
  MVD_WEIGHTS_VIRTUAL = 3      /* w_v / max(sum, 1) evaluated per voxel (NormalizingRandomAccess) */
};

/* PSF extraction from bead locations, on the GPU: ExtractPSF.extractNextImg
 * (deconv/ExtractPSF.java:277-296) = extractPSFLocal (:374-415: the image,
 * extended periodically, interpolated trilinearly at every PSF voxel around
 * every bead and summed in bead order) + normalize (:298-317: to [0, 1]) +
 * transformPSF (:325-372: resampled by the view's model into an odd box whose
 * centre is the transformed centre, zero out of bounds).
 * img: nx*ny*nz floats (the view as acquired); locations: nloc x (x, y, z)
 * doubles in pixel units; psf_size: (x, y, z).  psf_out receives the
 * normalised PSF in input calibration (getInputCalibrationPSFs);
 * transformed_out (nullable) the transformed PSF (getTransformedPSF), of the
 * size mvd_psf_transform_size returns; it needs the forward model and its
 * inverse (row-packed 3x4: AffineTransform3D.getRowPackedCopy() of the model
 * and of model.inverse()). */
int mvd_extract_psf(int device, const float* img, int nx, int ny, int nz, const double* locations, int nloc,
                    const int* psf_size, const double* model, const double* inverse_model, float* psf_out,
                    float* transformed_out);
/* Size (x, y, z) of the transformed PSF for a PSF of psf_size and the forward
 * model (ExtractPSF.transformPSF, deconv/ExtractPSF.java:325-372). */
int mvd_psf_transform_size(const int* psf_size, const double* model, int* new_size);

/* FFT shape of each block window (call before mvd_prepare). */
enum {
  MVD_FFT_SMOOTH = 0,   /* smallest 5-smooth size per axis (default) */
  MVD_FFT_POW2 = 1,     /* smallest power of two per axis (up to 1024) */
  MVD_FFT_AUTOTUNE = 2  /* time the smallest 5-smooth size and the smallest size of the register-resident
                           kernels (64, 128, 144, 192, 256, 288, 384, 512; twice those along x) per axis
                           on the GPU, keep the fastest shape (cached per window shape) */
};

/* new MVDeconInput + image geometry: nviews views of nx*ny*nz voxels on `device`. */
int mvd_create(int device, int nx, int ny, int nz, int nviews, mvd_session** out);
int mvd_destroy(mvd_session* s);

/* new MVDeconFFT(image, weight, kernel, ...) + MVDeconInput.add
 * (deconv/MVDeconFFT.java:79-172, deconv/MVDeconInput.java:48-54).
 * Views are added in order; weight may be NULL (then w = 1).  Kernel sizes must
 * be odd (cuda/BlockGeneratorFixedSizePrecise.java:43). */
int mvd_add_view(mvd_session* s, const float* img, const float* weight, const float* psf, int kx, int ky, int kz);

/* Input preparation of one view, on the GPU: the loop body of
 * ProcessForDeconvolution (deconv/ProcessForDeconvolution.java:187-277) with
 * TransformInputAndWeights (deconv/TransformInputAndWeights.java:100-135) or
 * TransformInput (deconv/TransformInput.java:69-95), then MVDeconFFT + add as
 * mvd_add_view.  raw: the view as acquired, rx*ry*rz floats, x-fastest.
 * inverse_affine: 12 doubles, the row-packed 3x4 inverse of the view's
 * registration (AffineTransform3D.inverse().getRowPackedCopy()), mapping
 * bounding-box coordinates to raw-view coordinates.  offset: the bounding
 * box's min (x, y, z); the session's image size is the bounding box's size.
 * The view becomes max(minValue, trilinear interpolation over the
 * mirror-single extension) where the mapped position lies inside the raw view
 * and 0 elsewhere.  border and range (x, y, z; both or neither): the view's
 * weight becomes its blending weight (BlendingRealRandomAccess.java:91-121)
 * at the mapped position -- PRECOMPUTED weights, normalised later by
 * mvd_set_weights; NULL: no weight (w = 1). */
int mvd_add_view_transformed(mvd_session* s, const float* raw, int rx, int ry, int rz, const double* inverse_affine,
                             const long long* offset, const float* border, const float* range, const float* psf,
                             int kx, int ky, int kz);

/* The same resampling without a session (host arrays in and out), for callers
 * that replace TransformInputAndWeights / TransformInput / TransformWeights
 * (deconv/TransformWeights.java:53-91) one view at a time.  out_img (nullable;
 * raw may then be NULL) receives nx*ny*nz floats as above; out_weight
 * (nullable) the blending weight, which needs border and range. */
int mvd_transform_view(int device, const float* raw, int rx, int ry, int rz, const double* inverse_affine,
                       const long long* offset, int nx, int ny, int nz, const float* border, const float* range,
                       float* out_img, float* out_weight);

/* mvd_transform_view on device memory: raw, out_img and out_weight are device
 * pointers on `device`, the launch is queued on `stream` (a hipStream_t; NULL =
 * the default stream) and the call returns without waiting.  No counterpart in
 * the reference, whose views are host images: for callers that keep the views
 * in HBM (a resampling step fused into a GPU pipeline). */
int mvd_transform_view_device(int device, const float* raw, int rx, int ry, int rz, const double* inverse_affine,
                              const long long* offset, int nx, int ny, int nz, const float* border,
                              const float* range, float* out_img, float* out_weight, void* stream);

/* Weight handling before preparation (ProcessForDeconvolution.java:325-405):
 * mode MVD_WEIGHTS_*; osem_index as EfficientBayesianBased.osemspeedupIndex
 * (deconv/EfficientBayesianBased.java:1023-1042, ProcessForDeconvolution.java:353-356):
 * 0 = 1, 1 = minimal number of overlapping views, 2 = average number,
 * 3 = osem_value.  norm_threads is the reference's thread count (the
 * average is the mean over its 2*norm_threads portions).
 * min_avg_out (nullable) receives {min overlapping views, avg overlapping views}. */
int mvd_set_weights(mvd_session* s, int mode, int osem_index, double osem_value, int norm_threads,
                    double* min_avg_out);

/* Choose the FFT shape policy (MVD_FFT_*); no counterpart in the reference,
 * whose FFT sizes are chosen by its FFT library. */
int mvd_set_fft_policy(mvd_session* s, int policy);

/* MVDeconInput.init(PSFTYPE) (deconv/MVDeconInput.java:62-68 -> MVDeconFFT.init,
 * deconv/MVDeconFFT.java:183-324) plus block split.  block_xyz NULL = the whole
 * image as one block of img+k-1 (MVDeconFFT.java:144-165); otherwise
 * BlockGeneratorFixedSizePrecise with that block size.  norm_threads is the
 * reference's Threads.numThreads(): AdjustInput.normImg counts the first of
 * its 2*norm_threads portions twice (deconv/AdjustInput.java:93-97). */
int mvd_prepare(mvd_session* s, int psf_type, double lambda, const int* block_xyz, int norm_threads);

/* Initial estimate (MVDeconvolution.java:116-148): psi0 == NULL fuses the
 * first iteration (average of views with data, 0.5 if NaN); otherwise psi0 is
 * used as given.  avg_out (nullable) receives the average. */
int mvd_init_psi(mvd_session* s, const float* psi0, double* avg_out);

/* n calls of MVDeconvolution.runIteration (MVDeconvolution.java:349-465).
 * stats (nullable) receives n * nviews * {sum change, max change}. */
int mvd_iterate(mvd_session* s, int n, double* stats);

/* Copy the estimate out; masked != 0 applies the final mask of
 * MVDeconvolution.java:201-208 (0 where no view has data) to the copy. */
int mvd_get_psi(mvd_session* s, float* out, int masked);

/* MVDeconFFT.getKernel1/getKernel2 (MVDeconFFT.java:373-374) of a view. */
int mvd_get_kernels(mvd_session* s, int view, float* kernel1, float* kernel2);

/* MVDeconFFT.setKernel (MVDeconFFT.java:364-369): replace a view's PSF and
 * re-run its init (normalisation, kernel2, spectra). */
int mvd_set_psf(mvd_session* s, int view, const float* psf);

/* setKernel for every view at once (MVDeconFFT.java:364-369 applied to all
 * views, then each view's init once): psfs[v] is view v's new PSF.  The state
 * equals calling mvd_set_psf for all views twice in a row -- the compound
 * kernels of a view read the other views' kernel1, so a single pass of
 * setKernel leaves the earlier views prepared from older PSFs -- at half the
 * preparation work. */
int mvd_set_psfs(mvd_session* s, const float* const* psfs);

/* Sum of the current (unmasked) estimate in double precision (no counterpart
 * in the reference: the total S of the PSF refinement step, computed where
 * the estimate lives instead of after a copy to the host). */
int mvd_psi_sum(mvd_session* s, double* out);

/* Poisson likelihood of the multiview forward model and its gradient (no
 * counterpart in the reference; its forward model is convolve1,
 * MVDeconFFT.java:384-470):  b_v = E_mirror(psi) * P_v,
 *   L = sum_v sum_x w_v (b_v - y_v ln b_v),  r_v = w_v (1 - y_v / b_v),
 *   dL/dP_v(j) = sum_x r_v(x) E(psi)(x + k/2 - j),  dL/dpsi = sum_v adjoint_v(r_v)
 * with the adjoint folding the mirror-single extension back.  psi == NULL uses
 * the session estimate; psfs == NULL (or psfs[v] == NULL) uses the views' PSFs
 * as given (not normalised).  grad_psi (n voxels) and grad_psf[v] (the view's
 * kernel box) are optional.  L and the psi gradient accumulate in FP64.
 * Once the session is prepared the gradient runs over the deconvolution's own
 * blocks (forward model and residual on every block first, then per block the
 * window spectrum is recomputed rather than kept: checkpointing, one block's
 * spectra in memory at a time); before that, over the single window
 * img + K - 1 (FFT lengths up to 1024). */
int mvd_nll_grad(mvd_session* s, const float* psi, const float* const* psfs, double* nll, float* grad_psi,
                 float* const* grad_psf);
/* As mvd_nll_grad, but dL/dpsi (when want_grad_psi) stays in device memory
 * until mvd_get_grad_psi copies it out (no counterpart in the reference: for
 * callers that keep the estimate on the GPU). */
int mvd_nll_grad_device(mvd_session* s, const float* psi, const float* const* psfs, double* nll, int want_grad_psi,
                        float* const* grad_psf);
int mvd_get_grad_psi(mvd_session* s, float* out);

/* Per-kernel-class GPU timing (HIP events on the session stream).
 * mvd_set_timing(s, 1) enables and resets; mvd_get_timing fills up to cap
 * classes: total ms and launch counts; returns the number of classes.
 * mvd_timing_name(i) names class i. */
int mvd_set_timing(mvd_session* s, int on);
int mvd_get_timing(mvd_session* s, double* total_ms, long long* launches, int cap);
const char* mvd_timing_name(int cls);

/* Geometry summary after mvd_prepare for one view: number of blocks and the
 * FFT size used; out[0..3] = {nblocks, qx, qy, qz}. */
int mvd_get_geometry(mvd_session* s, int view, int* out);

/* ---- Batched iteration over many stacks (no counterpart in the reference) --
 * A batch runs the iterations of several prepared sessions (independent
 * stacks of any size and view count) together: in view slot v every stack
 * with more than v views is updated, each FFT pass is one launch per kernel
 * class for all (stack, block) items, and with use_graph != 0 a whole
 * iteration is captured once and replayed as a hipGraph.  Per stack the
 * result is identical to mvd_iterate.  Sessions are borrowed: they must stay
 * alive and must not be iterated on their own while in a prepared batch.
 * stats (nullable): n iterations x (sum over stacks of nviews) x {sum, max},
 * stacks in the order they were added. */
typedef struct mvd_batch mvd_batch;
int mvd_batch_create(int device, mvd_batch** out);
int mvd_batch_add(mvd_batch* b, mvd_session* s);
int mvd_batch_prepare(mvd_batch* b, int use_graph);
int mvd_batch_iterate(mvd_batch* b, int n, double* stats);
int mvd_batch_destroy(mvd_batch* b);

/* ---- Output path of the deconvolved image ---------------------------------
 * EfficientBayesianBased.java:294-299 hands the result to an ImgExport with
 * the display range (0, 1); the exporters compute a missing range with
 * FusionHelper.minMax and the HDF5 export converts to 16 bit.  A TIFF file is
 * the counterpart of Save3dTIFF.exportImage (Save3dTIFF.java:79-133, ImageJ
 * FileSaver.saveAsTiffStack): one uncompressed page per z slice, an ImageJ
 * description (images, slices, min, max, spacing and origin when calibrated),
 * big-endian unless big_endian == 0; BigTIFF when the file needs 64-bit
 * offsets or bigtiff != 0.  Device inputs must be complete (synchronised)
 * when passed. */

/* FusionHelper.minMax (FusionHelper.java:108-172): minmax[0..1] = min, max
 * with java.lang.Math.min/max semantics (NaN propagates, -0.0 < 0.0), from
 * +/-Float.MAX_VALUE.  img: n floats in host memory (copied) or, for the
 * _device form, in device memory of `device` (stream: hipStream_t or NULL). */
int mvd_min_max(int device, const float* img, long long n, float* minmax);
int mvd_min_max_device(int device, const float* img, long long n, float* minmax, void* stream);

/* ExportSpimData2HDF5.updateAndGetMinMax (ExportSpimData2HDF5.java:227-253):
 * out = {max(0, min - (min + max) / 2 * 0.1), max + (min' + max) / 2 * 0.1}
 * with min' the widened min. */
int mvd_u16_range(const float* minmax, double* out);

/* RealUnsignedShortConverter(min, max) (imglib2 6.1.0, used at
 * ExportSpimData2HDF5.java:256-262) on n floats in device memory: out[i] =
 * min(65535, (int)(max(0, (img[i] - min) / (max - min) * 65535) + 0.5)),
 * NaN -> 0; byte-swapped when big_endian != 0.  Asynchronous on stream. */
int mvd_convert_u16_device(int device, const float* img, long long n, double min, double max, int big_endian,
                           unsigned short* out, void* stream);

/* Write a [nz][ny][nx] float image as a TIFF stack.  bits 32: float pages,
 * display range (min, max), NaN -> minMax; bits 16: the 16-bit conversion
 * over (min, max), NaN -> the widened minMax.  calibration (nullable) =
 * {xorigin, yorigin, zorigin, pixel size} as Save3dTIFF.java:103-108 sets it
 * from the bounding box.  img_on_device: img is device memory of `device`. */
int mvd_write_tiff(int device, const float* img, int img_on_device, int nx, int ny, int nz, int bits, double min,
                   double max, const double* calibration, int big_endian, int bigtiff, const char* path);

/* exporter.exportImage(getPsi(), ...) of a session: the masked estimate
 * written from device memory as by mvd_write_tiff. */
int mvd_export_psi(mvd_session* s, const char* path, int bits, double min, double max, const double* calibration,
                   int big_endian, int bigtiff);

/* Thread-local message of the last failed call. */
const char* mvd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MVDECON_H */

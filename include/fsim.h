/*
 * libfsim -- MI355X (gfx950) implementation of fairSIM's 2D SR-SIM
 * reconstruction path behind the org.fairsim.linalg vector API.
 *
 * C ABI only: plain pointers, sizes and opaque handles.  Every entry point
 * returns FSIM_OK (0) or a negative status; fsim_last_error() returns the
 * calling thread's last message.  The JNI shim maps FSIM_ERR_OOM to
 * OutOfMemoryError and every other failure to RuntimeException, which is the
 * reference's error convention (accel/cuda_common.h:33-63,
 * org/fairsim/accel/AccelVectorCplx.java:45-54).
 *
 * Vectors are device resident, row-major [h][w], complex interleaved
 * (re, im) float32 -- the layout of Vec2d (org/fairsim/linalg/Vec2d.java:23-24,
 * org/fairsim/linalg/BasicVectors.java:233-245).  All operations are
 * enqueued on one library stream in call order; results read back by
 * fsim_vec_download / the reductions are synchronous.
 */
#ifndef FSIM_H
#define FSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSIM_OK 0
#define FSIM_ERR_ARG (-1)     /* size mismatch / bad argument -> RuntimeException */
#define FSIM_ERR_OOM (-2)     /* device allocation failed -> OutOfMemoryError */
#define FSIM_ERR_HIP (-3)     /* HIP runtime error -> RuntimeException */
#define FSIM_ERR_STATE (-4)   /* not initialised / unsupported -> RuntimeException */
#define FSIM_ERR_BUSY (-5)    /* server: request refused (queue or memory full) */

typedef struct fsim_vec_s *fsim_vec;
typedef struct fsim_otf_s *fsim_otf;

/* Thread-local message of the last failing call on this thread. */
const char *fsim_last_error(void);

/* ---- lifecycle ---------------------------------------------------------
 * fsim_init: select the device and create the library stream.
 *   replaces AccelVectorFactory's JNI_OnLoad / factory set-up
 *   (accel/cuda_mgmt.cu:38-60, org/fairsim/accel/AccelVectorFactory.java:40-55)
 * fsim_sync: AccelVectorFactory.nativeSync (AccelVectorFactory.java:139;
 *   accel/cuda_mgmt.cu:38-44) */
int fsim_init(int device);
int fsim_sync(void);
/* Bytes of device memory currently held by vectors, raw buffers, channels
 * (tables and work buffers). */
int64_t fsim_device_bytes(void);
/* Test hook, no counterpart in the reference: the nth (0-based) channel or
 * work-buffer device allocation from now on fails as if the device were out
 * of memory (FSIM_ERR_OOM); nth < 0 switches the hook off.  For tests of the
 * error paths (every failure frees what the call had allocated). */
int fsim_debug_fail_alloc(int nth);

/* Raw buffers: AccelVectorFactory.nativeAllocMemory / nativeAllocMemoryHost
 * (AccelVectorFactory.java:141-142; accel/cuda_mgmt.cu:46-60).  Host buffers
 * are pinned.  fsim_memcpy copies on the library stream and returns when the
 * copy is done; kind: 1 host->device, 2 device->host, 3 device->device. */
int fsim_alloc_device(int64_t bytes, void **out);
int fsim_free_device(void *p);
int fsim_alloc_host(int64_t bytes, void **out);
int fsim_free_host(void *p);
int fsim_memcpy(void *dst, const void *src, int64_t bytes, int kind);

/* ---- vectors: AccelVectorReal / AccelVectorCplx natives --------------------
 * create: alloc(vf, n) (AccelVectorCplx.java:296, AccelVectorReal.java:218;
 *   accel/cuda_cplx.cu:39-70), zero-initialised.
 * destroy: dealloc (AccelVectorCplx.java:297; accel/cuda_cplx.cu:74-90).
 * upload/download: copyBuffer(J[FJZI)V (AccelVectorCplx.java:298;
 *   accel/cuda_cplx.cu:93-177); host layout = vectorData() float[]. */
int fsim_vec_create(int is_cplx, int w, int h, fsim_vec *out);
int fsim_vec_destroy(fsim_vec v);
int fsim_vec_shape(fsim_vec v, int *is_cplx, int *w, int *h);
/* Reinterpret the 2-D shape (w*h must equal the element count): the Java side
 * allocates by element count (alloc(vf, n)) and passes w, h with 2-D ops. */
int fsim_vec_reshape(fsim_vec v, int w, int h);
void *fsim_vec_device_ptr(fsim_vec v);
int fsim_vec_upload(fsim_vec v, const float *host);
int fsim_vec_download(fsim_vec v, float *host);

/* In-place arithmetic, `dst` is `this` of the Java method.
 * zero:      nativeZero (AccelVectorCplx.java:309; cuda_cplx.cu:331-335)
 * copy:      nativeCOPYCPLX / nativeCOPYREAL / Vec.Real.copy(Cplx)
 *            (AccelVectorCplx.java:306-307; AbstractVectorReal.java:99-106):
 *            cplx<-cplx, cplx<-real (im = 0), real<-cplx (real part), real<-real
 * add:       nativeAdd (AccelVectorCplx.java:300; cuda_cplx.cu:263-276)
 * add_const: nativeADDCONST (AccelVectorCplx.java:301; cuda_cplx.cu:279-288)
 * axpy:      nativeAXPY (AccelVectorCplx.java:302; cuda_cplx.cu:291-304),
 *            dst += (are + i aim) * x
 * times:     nativeTIMES / nativeTIMESREAL (AccelVectorCplx.java:303-304;
 *            cuda_cplx.cu:307-329); x complex or real; conj conjugates x
 * scal:      nativeScal (AccelVectorCplx.java:310; cuda_cplx.cu:338-341)
 * norm2:     nativeREDUCE(..., true) (AccelVectorCplx.java:305;
 *            cuda_cplx.cu:344-364): sum of float squares, double accumulation
 * sum:       AbstractVectorCplx.sumElements (AbstractVectorCplx.java:353-362),
 *            absent from the reference accelerator (it copies to the host) */
int fsim_vec_zero(fsim_vec v);
int fsim_vec_copy(fsim_vec dst, fsim_vec src);
int fsim_vec_add(fsim_vec dst, fsim_vec src);
int fsim_vec_add_const(fsim_vec dst, float re, float im);
int fsim_vec_axpy(fsim_vec dst, float are, float aim, fsim_vec x);
int fsim_vec_times(fsim_vec dst, fsim_vec x, int conj);
int fsim_vec_scal(fsim_vec dst, float re, float im);
int fsim_vec_norm2(fsim_vec v, double *out);
int fsim_vec_sum(fsim_vec v, double *re, double *im);

/* Vec2d ops.
 * set_u16:       nativeCOPYSHORT (AccelVectorCplx.java:483;
 *                cuda_cplx.cu:217-259, 375-381): out = (float)(in & 0xFFFF)
 *                (+0i); `host` points to w*h uint16 in host memory.
 * set_u16_device: same with the pixels already in device memory.
 * fft2d:         nativeFFT (AccelVectorCplx.java:487; cuda_cplx.cu:475-496);
 *                w, h powers of two in [2, 4096]; forward unscaled, inverse
 *                scaled by 1/(w*h).
 * fourier_shift: nativeFourierShift (AccelVectorCplx.java:488;
 *                cuda_cplx.cu:517-543), square vectors; the phase is rounded
 *                to float as BasicVectors.java:320 does, sin/cos use the
 *                MTool.fsin/fcos polynomial (linalg/MTool.java:34-59).
 * paste_freq:    nativePasteFreq (AccelVectorCplx.java:489;
 *                cuda_cplx.cu:547-580; BasicVectors.java:259-285).
 * get/set:       nativeGet / nativeSet (AccelVectorCplx.java:485-486). */
int fsim_vec_set_u16(fsim_vec v, const uint16_t *host);
int fsim_vec_set_u16_device(fsim_vec v, const uint16_t *dev);
int fsim_vec_fft2d(fsim_vec v, int inverse);
int fsim_vec_fourier_shift(fsim_vec v, double kx, double ky);
int fsim_vec_paste_freq(fsim_vec out, fsim_vec in, int xoff, int yoff);
int fsim_vec_get(fsim_vec v, int x, int y, float *re, float *im);
int fsim_vec_set(fsim_vec v, int x, int y, float re, float im);

/* ---- OTF / filters (sim_algorithm/OtfProvider.java, WienerFilter.java) ----
 * An fsim_otf holds OtfProvider's per-band 1-D tables on the device:
 * vals, valsAtt (complex) and valsOnlyAtt (real), nsamples each
 * (OtfProvider.java:39-41), plus cutoff and the table's cycles/micron per
 * sample (OtfProvider.java:44-45).
 *
 * fsim_otf_from_estimate: OtfProvider.fromEstimate + setAttenuation
 *   (OtfProvider.java:96-127, 276-287) computed on the host in double,
 *   writing nband=1 tables into caller arrays (vals/vals_att: 2*n floats,
 *   only_att: n floats) and cutoff/cpm.
 * fsim_otf_create: upload such tables (nband * nsamples); every value must be
 *   finite (FSIM_ERR_ARG otherwise). */
int fsim_otf_from_estimate(double na, double lambda_nm, double a, int nsamples,
                           double att_str, double att_fwhm, float *vals,
                           float *vals_att, float *only_att, double *cutoff,
                           double *cpm);
/* fsim_otf_attenuation: OtfProvider.setAttenuation's per-band tables
 *   (OtfProvider.java:276-287) for OTFs read from data: only_att[b][v] =
 *   valAttenuation(v * cpm, strength, fwhm), vals_att = vals * only_att
 *   (nband x nsamples, complex interleaved). */
int fsim_otf_attenuation(int nband, int nsamples, double cpm, double att_str, double att_fwhm,
                         const float *vals, float *vals_att, float *only_att);
int fsim_otf_create(int nband, int nsamples, double cutoff, double cpm,
                    const float *vals, const float *vals_att, const float *only_att,
                    fsim_otf *out);
int fsim_otf_destroy(fsim_otf o);

/* otfToVector (OtfProvider.java:331-373): write (write != 0) or multiply by
 * the conjugate (write == 0) the OTF of `band` centred at (kx, ky);
 * vec_cpm = the vector's cycles/micron per pixel (setPixelSize). */
int fsim_otf_to_vector(fsim_vec v, fsim_otf o, int band, double kx, double ky,
                       double vec_cpm, int use_att, int write);
/* maskOtf (OtfProvider.java:538-559) */
int fsim_mask_otf(fsim_vec v, fsim_otf o, double kx, double ky, double vec_cpm);
/* writeApoVector, real or complex vector (OtfProvider.java:437-492) */
int fsim_write_apo(fsim_vec v, fsim_otf o, double vec_cpm, double bend, double cutoff_fac);
/* writeAttenuationVector (OtfProvider.java:503-529) */
int fsim_write_attenuation(fsim_vec v, double vec_cpm, double str, double fwhm,
                           double kx, double ky);
/* WienerFilter(sp).getDenominator(w) (WienerFilter.java:46-122) into a real
 * vector: 1 / (sum_d sum_b |OTF_b(r1)|^2 att(r1) + |OTF_b(r2)|^2 att(r2) + w^2).
 * px, py: band-1 shift per direction (SimParam.Dir pX, pY). */
int fsim_wiener_denominator(fsim_vec out, fsim_otf o, double vec_cpm, int ndir,
                            int nband, const double *px, const double *py,
                            int use_att, double wparam);
/* SimUtils.fadeBorderCos (SimUtils.java:162-185) on a real vector. */
int fsim_fade_border(fsim_vec v, int px);

/* BandSeparation.createSeparationMatrix (BandSeparation.java:40-114) on the
 * host: phases[nphase] absolute band-1 phases, fac[nband] or NULL;
 * out: (2*nband-1) x nphase complex doubles, interleaved. */
int fsim_separation_matrix(const double *phases, int nphase, int nband,
                           const double *fac, double *out);
/* BandSeparation.separateBands multiply step (BandSeparation.java:152-163):
 * out[c] = sum_p S[c][p] * in[p], S given as (ncomp x nphase) complex floats
 * (the matrix after toFlt()).  One pass over the inputs. */
int fsim_separate_bands(const fsim_vec *in, int nphase, const fsim_vec *out, int ncomp,
                        const float *S);

/* ---- estimation (sim_algorithm/Correlation.java) ------------------------
 * locate_peak: Correlation.locatePeak (Correlation.java:144-167);
 *   res = {xPos, yPos, max, phase} in the reference's coordinates. */
int fsim_locate_peak(fsim_vec v, double kmin, double *res);

/* common_region: Correlation.commonRegion (Correlation.java:227-288), in
 *   place on both complex vectors.  The OTF weights (writeOtfVector, no
 *   attenuation) of bands bn0/bn1 centred at 0, (kx, ky) and (-kx, -ky) at the
 *   vector pixel size vec_cpm (cycles/micron per pixel); components below
 *   weight_limit are zeroed, the others divided by the weight when
 *   divide_by_otf; the DC ring ratio < dist or > 1 - dist is zeroed in band0
 *   and, shifted by ((int)kx, (int)ky), in band1. */
int fsim_common_region(fsim_vec band0, fsim_vec band1, int bn0, int bn1, fsim_otf otf, double kx, double ky,
                       double dist, double weight_limit, int divide_by_otf, double vec_cpm);
/* fit_peak: Correlation.fitPeak (Correlation.java:43-136).  Three rounds of
 *   a 10 x 10 search (+- search around the guess, search / 5 per round) for
 *   the shift of band1 against band0 (bands in Fourier space, unchanged);
 *   res = {kx, ky, phase, magnitude}; cntrl: NULL or a 30 x 10 float image
 *   (row-major, width 30) of the normalised correlation magnitudes. */
int fsim_fit_peak(fsim_vec band0, fsim_vec band1, int bn0, int bn1, fsim_otf otf, double kx, double ky,
                  double weight_limit, double search, double vec_cpm, float *cntrl, double *res);
/* get_peak: Correlation.getPeak (Correlation.java:183-209): the
 *   cross-correlation of band0 and band1 shifted to (kx, ky) over their common
 *   region, scaled by 1/|band0|^2; res = {re, im} (Cplx.Double). */
int fsim_get_peak(fsim_vec band0, fsim_vec band1, int bn0, int bn1, fsim_otf otf, double kx, double ky,
                  double weight_limit, double vec_cpm, double *res);

/* ---- batched parameter estimation ------------------------------------------
 * SimAlgorithm.estimateParameters (SimAlgorithm.java:37-256) for `count`
 * stacks at once, with the input preparation of the GUI
 * (sim_gui/ImageControl.java:816-840: fadeBorderCos, copy to a complex
 * vector, forward FFT).  Per stack and direction: band separation with
 * equidistant phases 2 pi p / nphase and default factors, the coarse
 * correlation and locatePeak (exclusion fit_exclude * cutoff / vec_cpm),
 * fitPeak(band 0, fit band, -peak, weight limit 0.05, search 2.5) and getPeak
 * of band 1 (and band 2 for three bands) — the same arithmetic as the
 * per-call entry points above, in about 40 stream-ordered launches for the
 * whole batch and one synchronisation. */
typedef struct {
    int n;              /* image side; square, a power of two              */
    int ndir;           /* pattern directions                              */
    int nphase;         /* phases per direction (equidistant)              */
    int nband;          /* 2 or 3                                          */
    int fit_band;       /* 1 or 2                                          */
    double fit_exclude; /* locatePeak exclusion, fraction of the OTF cutoff */
    double vec_cpm;     /* cycles per micron of the n x n vector           */
    int fade_px;        /* fadeBorderCos border applied to every image; 0: none */
    int live;           /* 0: SimAlgorithm.estimateParameters (SimAlgorithm.java:37-256);
                         * 1: the live-mode refit ReconstructionRunner.ParameterRefitThread.
                         * doRefit (livemode/ReconstructionRunner.java:642-746): the peak is
                         * searched on the highest band's correlation (fit_band ignored;
                         * the reference uses fit_exclude = 0.8), fitPeak correlates it with
                         * OTF band numbers (0, 1) whatever the band, getPeak of band 1 at
                         * peak / (bands - 1) gives phase and modulation 1, modulation 2 is
                         * not estimated (mod2 = NaN: fsim_params_from_estimate keeps the
                         * base value); the input is faded by multiplying with the faded
                         * vector of ones (times(borderDampen), :614-617, 661-664) instead of
                         * fadeBorderCos on each image */
} fsim_estimate_params;

typedef struct {
    double px, py;          /* shift as SimParam.Dir.setPxPy receives it (-peak)  */
    double pha_off;         /* phase of getPeak on band 1 (setPhaOff)             */
    double mod1, mod2;      /* |getPeak| of bands 1 and 2 (mod2 = 0 with 2 bands,
                             * NaN when not estimated: the live refit)             */
    double peak_x, peak_y;  /* coarse correlation peak (locatePeak convention)    */
} fsim_estimate_result;

/* estimate: images = device pointer (16-byte aligned) to count x ndir x
 *   nphase x n x n float32 (direction-major, as the stacks are taken);
 *   res = host array of
 *   count x ndir results.  Errors: FSIM_ERR_ARG for unsupported sizes or
 *   bands (bands beyond the OTF's), or a shift larger than the vector in
 *   commonRegion (the reference fails there with an index exception). */
int fsim_estimate(const float *images, int count, const fsim_estimate_params *p, fsim_otf otf,
                  fsim_estimate_result *res);
/* Test hook, no counterpart in the reference: on != 0 makes fsim_estimate's
 * commonRegion pass form the shifted OTF weight at every pixel instead of
 * classifying pixels by squared distance through per-table run starts; the
 * two give the same bits.  For tests that compare them. */
int fsim_debug_region_exact(int on);

/* ---- fused fast-shift reconstruction --------------------------------------
 * One reconstruction = livemode/ReconstructionRunner.java:456-542 /
 * tests/TestAccelSpeed.java:237-252, 563-703 for one channel: raw stack of
 * ndir*nphase frames n x n uint16 (angle-major) -> 2n x 2n float32.
 *
 * fsim_channel: per-channel resident state built once (the FilterUpdateThread
 * of ReconstructionRunner.java:773-828): OTF tables, reciprocal Wiener
 * denominator and apodization for the (2n)^2 grid, the fade-border factors,
 * and the per-direction separation matrices and shift parameters. */
typedef struct fsim_channel_s *fsim_channel;

typedef struct {
    int n;                    /* raw side, power of two in [64, 2048] */
    int ndir, nphase, nband;  /* ndir <= 5, nphase <= 7, nband <= 3 */
    double px[5], py[5];      /* band-1 shift per direction (SimParam.Dir pX, pY) */
    double phases[5][7];      /* absolute band-1 phases (SimParam.Dir.getPhases) */
    double modul[5][3];       /* raw modulations; clamped as getModulations */
    double vec_cpm;           /* 1 / (n * micronsPerPixel) */
    int use_att;              /* OTF with attenuation */
    double wiener;            /* Wiener parameter w */
    double apo_bend, apo_cut; /* writeApoVector(bend, cutoff factor) */
    int fade_px;              /* fadeBorderCos width (10), 0 = off */
    int mode;                 /* 0: derive each band's neg component from its pos
                                 component (exact for real raw data and a real
                                 OTF; chosen only when the OTF is real),
                                 1: compute every component separately */
    int otf_band0;            /* 1: band 0's OTF for every component, as the live
                                 ReconstructionRunner builds one OTF vector per
                                 channel (ReconstructionRunner.java:786-789);
                                 0: each band's own (tests/TestAccelSpeed.java:
                                 506-514).  Equal for single-band OTFs. */
} fsim_recon_params;

int fsim_channel_create(const fsim_recon_params *p, fsim_otf otf, fsim_channel *out);
int fsim_channel_destroy(fsim_channel ch);
/* The channel's size class: raw side, directions, phases, stored components
 * per direction (bands with the Hermitian pairing, 2*bands-1 without) and
 * whether the result is Hermitian (real OTF, mode 0).  No counterpart in the
 * reference; the key of fsim_batch_* classes and of fsim_autotune_*. */
int fsim_channel_class(fsim_channel ch, int *n, int *D, int *P, int *Cs, int *half);

/* Reconstruct `count` stacks of one channel.  raw_dev: count stacks back to
 * back in device memory; out_dev: count images of (2n)^2 float.  `stream` is
 * a hipStream_t or NULL for the library stream.  Asynchronous.  Work buffers
 * are held by the channel per stream and grow to the largest count seen on
 * that stream, so calls on different streams (one per thread, as the
 * reference's ReconstructionThreads each own their vectors,
 * livemode/ReconstructionRunner.java:243-251, 398-413) run concurrently;
 * calls on one stream from several threads must be serialised by the caller.
 * Replaces, fused, the per-frame op sequence of
 * ReconstructionRunner.ReconstructionThread.run (:415-594). */
int fsim_reconstruct(fsim_channel ch, const uint16_t *raw_dev, float *out_dev,
                     int count, void *stream);

/* reconstruct_noalloc: fsim_reconstruct with the work buffers (the A/B
 * intermediates, 10..16 * (2n)^2 bytes per stack) in a fixed device array of
 * the library instead of allocated ones.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid it).  The call allocates
 * nothing and frees nothing; the array (128 MiB per GPU, held until the
 * process ends) bounds count: a call whose work buffers exceed it fails with
 * FSIM_ERR_ARG before anything is launched (at 512^2, up to 6 stacks per
 * call; at 1024^2, 1; 2048^2 does not fit).  The array is shared by every
 * call on the GPU: calls that overlap in time (on different streams or
 * threads) race and must be serialised by the caller.  Nothing in it needs to
 * start at zero.  The channel must have been created (its tables and the
 * transform tables exist from then on).  Asynchronous on `stream`. */
int fsim_reconstruct_noalloc(fsim_channel ch, const uint16_t *raw_dev, float *out_dev,
                             int count, void *stream);

/* SIM parameters of one stack (SimParam.Dir per direction): band-1 shift
 * pX, pY; the phases used for the separation (getPhases: phases + phaOff);
 * modulations per band (clamped to [0.3, 1.1] as getModulations). */
typedef struct {
    double px[5], py[5];
    double phases[5][7];
    double modul[5][3];
} fsim_sim_params;

/* reconstruct_params: fsim_reconstruct with each stack's own SIM parameters
 *   (sp[count], host memory) -- the reconstruction that follows a per-stack
 *   estimate (SimAlgorithm.runReconstruction after estimateParameters,
 *   SimAlgorithm.java:268-709; the live refit, ReconstructionRunner.java:
 *   642-746).  The channel supplies the OTF, the Wiener constant,
 *   attenuation, apodization and fade; the Wiener denominator
 *   (WienerFilter.java:46-122), separation matrices and shift ramps are built
 *   on the device for every stack.  Asynchronous. */
int fsim_reconstruct_params(fsim_channel ch, const uint16_t *raw_dev, float *out_dev, int count,
                            const fsim_sim_params *sp, void *stream);

/* params_from_estimate: what estimateParameters leaves in the SimParam
 *   (SimAlgorithm.java:136-142, 155-158) as per-stack reconstruction
 *   parameters: for stack s and direction d, pX, pY = (px, py) / (nband - 1)
 *   (SimParam.Dir.setPxPy, SimParam.java:536-541), phases = base phases +
 *   pha_off (setPhaOff; getPhases, :472-477), modulation 1 (and 2) from the
 *   estimate, the others from base (also modulation 2 when the estimate
 *   holds NaN for it, as the live refit leaves it).  res: count x ndir results of
 *   fsim_estimate; base: the parameters without phase offset; out: count
 *   entries (host memory). */
int fsim_params_from_estimate(const fsim_estimate_result *res, int count, int ndir, int nphase, int nband,
                              const fsim_sim_params *base, fsim_sim_params *out);

/* ---- filter styles beyond Wiener ------------------------------------------
 * rl_deconvolve: RLDeconvolution.deconvolve(img, otf, steps,
 *   inputIsInFreqSpace) (sim_algorithm/RLDeconvolution.java:43-103), in
 *   place; img and otf complex vectors of one power-of-two size.  Used by
 *   SimAlgorithm.runReconstruction's "RL on input / output / both" filter
 *   styles (SimAlgorithm.java:320-356, 587-636).
 * clip_scale: SimUtils.clipAndScale (SimUtils.java:244-264) on a real
 *   vector: scale by 255 / (largest positive value), then clip below 0. */
int fsim_rl_deconvolve(fsim_vec img, fsim_vec otf, int steps, int input_in_freq);
int fsim_clip_scale(fsim_vec v, int clip, int scale);

/* ---- live-mode side outputs ---------------------------------------------
 * widefield: the widefield image of each of `count` stacks (n x n float):
 *   the faded raw frames (dampBorder, fadeBorderCos of ones) summed in
 *   (direction, phase) order, times 1.f / (ndir * nphase)
 *   (ReconstructionRunner.java:452-481); raw_output = k in 1..ndir*nphase
 *   gives raw frame k instead (the reference's rawOutput).  The camera offset
 *   of a channel (PerChannel.offset, :75, 92) is read by the reference but
 *   never applied, so there is nothing to apply here either.
 * registration: Registration (registration/Registration.java): per-pixel
 *   source positions for reconstruction-size images (width x height tables,
 *   the "X Trans" / "Y Trans" blocks of the registration file); widefield
 *   images use the tables at even positions, halved (:371-385).  Images
 *   smaller than the tables are centred in them.  register_images: bilinear
 *   transPixel (:239-268) for `count` images of w x h (dst != src); pixels
 *   whose source lies outside the image are 0. */
int fsim_widefield(fsim_channel ch, const uint16_t *raw_dev, float *out_dev, int count, int raw_output,
                   void *stream);
typedef struct fsim_registration_s *fsim_registration;
int fsim_registration_create(const float *xtrans, const float *ytrans, int width, int height,
                             fsim_registration *out);
int fsim_registration_destroy(fsim_registration r);
int fsim_register_images(fsim_registration r, int widefield, const float *src_dev, float *dst_dev, int w, int h,
                         int count, void *stream);

/* ---- request server (continuous batching) ---------------------------------
 * The serving form of livemode/ReconstructionRunner.java: submit() is
 * queueImage() (:140-150) -- it refuses with FSIM_ERR_BUSY when max_queue
 * requests are pending, as the reference drops frames when its queue is full
 * (:366-370) -- and wait() takes a result as finalRecon.take() does (:57-58).
 * A scheduler thread batches pending requests of one size class (n, directions,
 * phases, bands; channels may differ within a batch), admits a batch when
 * fewer than nslots batches are in flight and an idle device slot of that
 * class exists or mem_budget allows a new one (idle slots of other classes
 * are freed, least recently used first, only as far as the budget needs:
 * freeing device memory waits for the device), and runs each batch on its
 * slot's stream, so that up to nslots batches overlap on the device;
 * requests arriving while batches run join the next free slot.
 * add_channel allocates a slot for the channel's class when none exists and
 * the budget has room, so the first requests do not wait for it.  Channels stay owned by the caller and
 * must outlive the server.  Host buffers are copied asynchronously: pass
 * pinned memory (fsim_alloc_host) for overlap.  FSIM_SUBMIT_DEVICE: raw/out
 * are device pointers.  Results are bitwise independent of batch composition.
 * A request whose slot alone exceeds mem_budget is refused at submit
 * (FSIM_ERR_BUSY); one whose slot cannot be allocated while nothing else is
 * on the device fails (wait returns FSIM_ERR_OOM) instead of blocking the
 * queue. */
typedef struct fsim_server_s *fsim_server;
typedef struct {
    int64_t mem_budget;  /* device bytes for slots (work buffers, staging) */
    int max_batch;       /* stacks per launch */
    int max_queue;       /* pending requests before submit refuses */
    int nslots;          /* batches in flight at once (streams) */
} fsim_server_config;
typedef struct {
    int64_t submitted, completed, rejected, failed, batches, pending, device_bytes;
    double mean_batch, p50_ms, p99_ms, max_ms;  /* latency: submit -> result ready */
} fsim_server_stats;
#define FSIM_SUBMIT_DEVICE 1
int fsim_server_create(const fsim_server_config *cfg, fsim_server *out);
int fsim_server_destroy(fsim_server s);
int fsim_server_add_channel(fsim_server s, fsim_channel ch, int *id);
int fsim_server_submit(fsim_server s, int channel, const uint16_t *raw, float *out, int flags,
                       uint64_t *ticket);
/* submit_params: one request of `count` stacks (e.g. the timepoints of one
 * acquisition, raw = count consecutive stacks, out = count consecutive
 * images) reconstructed with its own SIM parameters *sp (copied) instead of
 * the channel's -- the live refit (ReconstructionRunner.java:642-746) or a
 * fit on the first timepoint applied to the following ones.  The request
 * gets a slot of its own, sized by its footprint (work buffers for count
 * stacks plus one set of parameter tables) and admitted against mem_budget
 * like the batches; a request larger than the budget is refused
 * (FSIM_ERR_BUSY).  Results are bitwise those of fsim_reconstruct_params
 * with sp repeated count times. */
int fsim_server_submit_params(fsim_server s, int channel, const uint16_t *raw, float *out, int count,
                              const fsim_sim_params *sp, int flags, uint64_t *ticket);
/* timeout_ms < 0 waits forever; FSIM_ERR_BUSY on timeout */
int fsim_server_wait(fsim_server s, uint64_t ticket, int timeout_ms, double *latency_ms);
int fsim_server_get_stats(fsim_server s, fsim_server_stats *st, int reset);

/* ---- fixed-shape batches: mixed sizes and channels, one graph launch -------
 * No single counterpart in the reference: it reconstructs image by image from
 * several threads (livemode/ReconstructionRunner.java:243-251, 380-413).
 * fsim_batch_create binds slot i to channels[i]; the slots are grouped into
 * size classes (raw side, directions, phases, stored components), each with
 * its own work buffers, device job list and stream.  fsim_batch_run
 * reconstructs raw[i] (device, D*P frames of n*n uint16) into out[i] (device,
 * (2n)^2 float32) for every slot, stream-ordered on `stream` (NULL: the
 * library stream): the class pipelines run concurrently, and with
 * flags & FSIM_BATCH_GRAPH the launch sequence is captured as a hipGraph on
 * the first run and replayed afterwards, the slot addresses being rewritten
 * in the device job lists before each replay.  With flags &
 * FSIM_BATCH_PERSISTENT the slots of sides 256 to 2048 run through
 * persistent kernels instead of per-class launches: for each pipeline stage
 * one launch per group of sides (256 and 512; 1024 and 2048) walks a device
 * list of (slot class, tile or row block) work items, larger side first;
 * other sides keep their class launches.  raw[i] and out[i] must be 8-byte
 * aligned.  One run at a time per batch (runs on one stream serialise).
 * Results are bitwise those of fsim_reconstruct on the same stacks with the
 * default launch choices (no autotuned choice set for the class). */
#define FSIM_BATCH_GRAPH 1
#define FSIM_BATCH_PERSISTENT 2
typedef struct fsim_batch_s *fsim_batch;
int fsim_batch_create(const fsim_channel *channels, int count, int flags, fsim_batch *out);
int fsim_batch_run(fsim_batch b, const uint16_t *const *raw, float *const *out, void *stream);
int fsim_batch_info(fsim_batch b, int *nclasses, int *captured, int64_t *device_bytes);
int fsim_batch_destroy(fsim_batch b);

/* ---- autotuner of the pipeline's launch shapes --------------------------
 * No counterpart in the reference (its cuFFT plans are fixed per size,
 * accel/FFTProvider.java:63-102).  Per size class (n, D, P, stored
 * components Cs, Hermitian result half) the pipeline chooses: sep = 1 the
 * separation with all units of a row in one wavefront (n <= 512, D*P <= 9),
 * 0 one unit per block; col = the column pass's candidate 0..7 (elements per
 * thread = Stockham radix 8, 16 or 32, columns per block 4, 8 or 16; the
 * shape is reported in col_elems / col_width, col_width negated for the
 * layouts with one column's transform per wavefront).  fsim_autotune times every candidate on
 * the channel's class (count zero-filled stacks; after a 30 ms warm-up, three
 * interleaved rounds of `reps` timed launches each, HIP events, best round
 * per candidate) and caches the fastest -- the default unless a candidate is
 * more than 2% faster -- for the process; every later launch of
 * that class (fsim_reconstruct, _params, the server, batches created after)
 * uses it.  save/load persist the cache as text ("fsim-autotune 2" then
 * "n D P Cs half sep col chunk streams count ms" per line; version 1 files
 * without the chunk columns load too); load merges and rejects a
 * malformed file whole.  set pins a choice by hand, get returns -1s when the
 * class has no entry.  These calls do not need a GPU except fsim_autotune.
 * fsim_init loads the file named by the environment variable
 * FSIM_AUTOTUNE_CACHE, when set (a missing or malformed file is ignored). */
typedef struct {
    int sep, col, col_elems, col_width, candidates;
    double ms_best, ms_default;   /* per call of `count` stacks: chosen, all defaults */
    int chunk, streams;           /* chosen chunking: stacks per chunk (0: one launch), streams */
    double ms_whole, ms_chunk_default;  /* chosen shapes as one launch / with the size's default chunking */
} fsim_tune_result;
int fsim_autotune(fsim_channel ch, int count, int reps, fsim_tune_result *res);
int fsim_autotune_set(int n, int D, int P, int Cs, int half, int sep, int col);
int fsim_autotune_get(int n, int D, int P, int Cs, int half, int *sep, int *col);
/* Chunked calls: fsim_reconstruct runs a call's stacks as consecutive chunks
 * of `chunk` stacks, round-robin on `streams` library streams forked from and
 * joined into the caller's stream, each stream reusing its own chunk-sized
 * work buffers (which then stay in the GPU's 256 MB Infinity Cache between
 * the passes, and one chunk's tail overlaps the next one's head).  chunk =
 * -1: the size default (up to 512^2: the largest power of two of stacks
 * whose work buffers fit 160 MiB, on 2 streams; one launch above), 0: one
 * launch.  Results are bitwise those of one launch.  fsim_autotune times the
 * chunkings after the shapes.  get_chunks reports the plan a call of `count`
 * stacks of the class uses (chunk = count, streams = 1 for one launch). */
int fsim_autotune_set_chunks(int n, int D, int P, int Cs, int half, int chunk, int streams);
int fsim_autotune_get_chunks(int n, int D, int P, int Cs, int half, int count, int *chunk, int *streams);
int fsim_autotune_clear(void);
int fsim_autotune_save(const char *path);
int fsim_autotune_load(const char *path);

/* ---- live stream ingest (transport/ImageWrapper.java,
 * livemode/SimSequenceExtractor.java) -----------------------------------------
 * Wire format: a 128-byte little-endian header per camera frame, then the
 * pixels (transport/header-definition.md; ImageWrapper.writeHeader /
 * parseHeader, ImageWrapper.java:420-476): seq_nr @0 (int64), pixel count
 * w*h @12 (int32), version @16 (1), bytes per pixel @17 (1 or 2), width @32,
 * height @34, posA @36 (angle), posB @38 (phase), pos0 @40 (z), pos1 @42
 * (channel), pos2 @44 (int32, time index), camera / capture / record time
 * stamps @48/56/64 (int64, microseconds).  ImageWrapper neither writes nor
 * checks the "LIVESIM" tag of the header description; neither does this. */
typedef struct {
    int64_t seq_nr;
    int32_t npix;                 /* width * height */
    int32_t version, bpp;
    int32_t width, height;        /* int16 on the wire */
    int32_t pos_a, pos_b, pos0, pos1;
    int32_t pos2;
    int64_t time_camera, time_capture, time_record;
} fsim_frame_header;

/* parse: ImageWrapper.parseHeader's checks (width/height beyond the maximum,
 *   w*h != count, bpp not 1 or 2, version not 1) fail with FSIM_ERR_ARG.
 * write: ImageWrapper.writeHeader (zero-filled, version 1). */
int fsim_wire_parse_header(const uint8_t *hdr, int max_width, int max_height, fsim_frame_header *out);
int fsim_wire_write_header(const fsim_frame_header *h, uint8_t *hdr);

/* Sequence assembly: SimSequenceExtractor's per-channel threads
 * (ImageSorter :146-178, PerChannelBuffer :211-470) as a push-driven state
 * machine.  Frames are routed by pos1 to the channel with that number, put
 * in sequence-number order through a sort buffer (:246-330; on overflow the
 * frame is dropped and the buffer resynchronised to the lowest of the next
 * ndir*nphase frames), a sync frame is searched (sync_mode 0: a camera-time
 * gap > sync_delay before a frame, which then starts the sequence; 1: a
 * camera-time gap within 50 of sync_delay, or a frame whose mean pixel
 * exceeds sync_avr -- the frame after a bright one is skipped), and then
 * seq_count sequences of ndir*nphase frames are taken (:369-470).  Frames of
 * 8-bit packets are sign-extended to 16 bit as ImageWrapper.getPixels does.
 * Assembled sequences wait in a per-channel queue of queue_capacity (further
 * ones are dropped and counted, :457-460). */
typedef struct fsim_ingest_s *fsim_ingest;
typedef struct {
    int ndir, nphase;
    int sync_mode;          /* 0 or 1 (SyncMode) */
    int64_t sync_delay;     /* SyncFrameDelay */
    int sync_avr;           /* SyncFrameAvr */
    int seq_count;          /* SyncFrameFreq: sequences per sync frame */
    int sort_capacity;      /* reference: 20 */
    int queue_capacity;     /* reference: (ndir*nphase)^2 * 10 */
    int max_width, max_height;
} fsim_ingest_config;
typedef struct {
    int64_t frames, sequences, missed_sequences, dropped_frames, sync_frames, no_sync_windows;
} fsim_ingest_stats;
typedef struct {
    int channel;            /* index into channel_numbers */
    int width, height;
    int64_t first_seq_nr, time_camera;
    int32_t pos2;           /* time index of the first frame */
} fsim_ingest_seq_info;

int fsim_ingest_create(const fsim_ingest_config *cfg, const int *channel_numbers, int nchannels, fsim_ingest *out);
int fsim_ingest_destroy(fsim_ingest in);
/* one packet (header + payload, len bytes); FSIM_ERR_ARG for a broken header,
 * a short packet or an unknown channel (the packet is dropped, as the
 * reference logs and drops it) */
int fsim_ingest_push(fsim_ingest in, const uint8_t *packet, int64_t len);
/* the next assembled sequence of channel index ch: 1 and the frames
 * (ndir*nphase x w x h uint16, `out` holds max_pixels of them) when one is
 * waiting, else 0 */
int fsim_ingest_pop(fsim_ingest in, int ch, uint16_t *out, int64_t max_pixels, fsim_ingest_seq_info *info);
int fsim_ingest_get_stats(fsim_ingest in, int ch, fsim_ingest_stats *st);

/* Feeding the request server directly: sequences of channel index c are
 * submitted to server channel server_channels[c] (frames must be n x n;
 * others are dropped and counted) through pinned staging of up to
 * max_inflight requests; push() and result() keep the pipe full.  result()
 * waits (timeout_ms < 0: forever; FSIM_ERR_BUSY on timeout) for the oldest
 * submitted sequence and copies its (2n)^2 image to out -- the
 * ReconstructionRunner's finalRecon.take() (ReconstructionRunner.java:57-58). */
int fsim_ingest_attach_server(fsim_ingest in, fsim_server srv, const int *server_channels, int n, int max_inflight);
int fsim_ingest_result(fsim_ingest in, int timeout_ms, float *out, fsim_ingest_seq_info *info);

/* Stage timing of the fused pipeline (no counterpart in the reference, which
 * only has Tool.Timer wall clocks, utils/Tool.java:210-245): while enabled,
 * every fsim_reconstruct records HIP events around its kernels on its stream;
 * fsim_stage_times waits for them and returns the summed milliseconds and
 * launch counts of stages 0..4 = band separation + row FFT, shifted column
 * pass, shifted row pass, gather + filter + column IFFT, row IFFT. */
int fsim_stage_timing(int enable);
int fsim_stage_times(double *ms, int *launches, int nstage);

#ifdef __cplusplus
}
#endif
#endif /* FSIM_H */

/*
 * gisapr_hip.h -- C ABI of libgisapr_hip.so, the gfx950 (MI355X) engine for
 * GisAPR's candidate-scoring path.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers (hipMalloc / torch CUDA tensors)
 *    unless stated; sizes are element counts.  Complex arrays are
 *    interleaved float32 (re, im) = complex64.
 *  - `stream` is a hipStream_t (NULL = default stream).  Every call is
 *    asynchronous on that stream; nothing here synchronises.
 *  - Return value: 0 on success, > 0 a hipError_t, < 0 an argument error
 *    (-1 invalid, -2 unsupported size).  gis_last_error() gives the message.
 *  - Box sizes compiled: N in {32, 64, 96, 128, 192, 256}.  The packed
 *    Fourier layout of an n x n image is the rfft2 layout [n][n/2+1] with
 *    rows in natural frequency order (v606:494-541).  Banks and particle
 *    spectra are stored UNshifted; the reference's fftshift "pre-shift"
 *    (v606:535-537) is folded into the correlation tables.
 *
 * Reference files cited: v606 = test1_with_isspa_weight_varingKK_search_
 * translation_also_v606_torch_optimized_standalone.py; p2m =
 * pdb2mrc_gpu_ver_fp32_v3.py; ZL = GisAPR-Lite-v1.3.2.zip/GisAPR-210-DE-v1.3.2.
 */
#ifndef GISAPR_HIP_H
#define GISAPR_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call on this thread. */
const char* gis_last_error(void);
/* ABI version (1). */
int gis_version(void);
/* Makes `device` current for the calling thread and looks up the device's
 * workspace of gis_correlate_split_noalloc. */
int gis_init(int device);

/*
 * (a) Atom -> voxel splat.  Replaces ZL:gisapr_cuda_density.py:21-22
 * deposit(xyz, numbers, volume, n, box, width, apix, rp, kn) and the
 * index_add_ loop of density_from_coords (p2m:41-98).  ADDS into `volume`
 * ([box][box][box] float32, [z][y][x]); the caller zeroes it or pre-fills it
 * with a resident fixed-body density.  xyz [n][3] float32 (A), numbers [n]
 * float32 atomic numbers (0 = skipped); width/rp/kn as p2m:53-58.
 */
int gis_splat(const float* xyz, const float* numbers, int n_atoms, int box, int width,
              double apix, double rp, double kn, float* volume, void* stream);

/*
 * (a) Rigid move + 8.3 round trip of a chain for a batch of poses.  Replaces
 * transform_coordinates (pdb_text.py:123-142: (x - centre) @ R + centre + t)
 * and the PDB write / re-read of the moved chain (Z3 pdb_text.py:179-198:
 * float(f"{v:8.3f}")) for each candidate pose.  xyz [n][3] float32, centre
 * 3 floats (host memory; the chain's float32 mean), rt [poses][12] float64
 * DEVICE array (R row-major, then t), out [poses][n][3] float32.  The float64
 * product is the fixed chain d0 R0j, then fma(d1, R1j, .), fma(d2, R2j, .)
 * (NumPy's float64 matmul of the same operands), the rounding to 3 decimals
 * is exact (ties to even on the exact value).  The caller checks that the
 * moved coordinates fit the 8.3 field.
 */
int gis_move_chain(const float* xyz, int n_atoms, const float* centre, const double* rt, int poses, float* out,
                   void* stream);

/*
 * (b, prefix) Pad / sinc^2 / fftshift / rfftn of the model volume.
 * Replaces IntegratedProjector.build_templates v606:1068-1073.
 * volume [N][N][N] float32 -> dft: the [2N][2N][N+1] complex64 half spectrum
 * in natural frequency order, stored in bricks of 2 (z) x 2 (y) x 4 (x)
 * values: element (kz, ky, kx) is at
 *   ((((kz/2) * N + ky/2) * XB + kx/4) * 2 + kz%2) * 2 + ky%2) * 4 + kx%4,
 * XB = N/4 + 1 (gis_volume_dft_bytes(N) bytes).  ws: gis_volume_dft_workspace(N)
 * bytes of scratch (the x / y intermediate of the N data planes,
 * N * N * ceil((N+1)/8) * 16 complex64 values).
 */
size_t gis_volume_dft_bytes(int N);
size_t gis_volume_dft_workspace(int N);
int gis_volume_dft(const float* volume, int N, void* dft, void* ws, size_t ws_bytes, void* stream);

/*
 * The same transform padded (pad 1, L = 2N, as above) or unpadded (pad 0,
 * L = N: project3d_and_whiten_cuda.py --skip_padding, :103-117): the
 * [L][L][L/2+1] half spectrum in bricks of 2 (z) x 2 (y) x 4 (x) values,
 *   ((((kz/2) * L/2 + ky/2) * XB + kx/4) * 2 + kz%2) * 2 + ky%2) * 4 + kx%4,
 * XB = L/8 + 1, with sinc^2 over fftshift(fftfreq(L)).
 */
size_t gis_volume_dft_bytes_ex(int N, int pad);
size_t gis_volume_dft_workspace_ex(int N, int pad);
int gis_volume_dft_ex(const float* volume, int N, int pad, void* dft, void* ws, size_t ws_bytes, void* stream);

/* Workspace bytes gis_build_templates needs for M directions at box N. */
size_t gis_build_templates_workspace(int N, int M);
/* The same for the padded (pad 1) or unpadded (pad 0) slice transform. */
size_t gis_build_templates_workspace_ex(int N, int M, int pad);
/* Template workgroups resident on the current device at once (box N, pad as
 * above); launches of fewer directions than this leave compute units idle.
 * No reference counterpart: sizes the batched builders' candidate groups. */
int gis_build_templates_resident(int N, int pad);

/*
 * (b) Template bank.  Replaces the per-direction loop of
 * IntegratedProjector.build_templates (v606:1077-1147): central-slice
 * gather with Friedel conjugation, irfft2, crop, whitening
 * (ProjectionWhitening batch==1 path v606:938-992), mask, fft2, crop to n,
 * pack, x FSC, unit L2 norm.
 *   dft  bricked half spectrum from gis_volume_dft
 *   rts  [M][3][3] float32: Euler(rot, tilt, 0)^T per direction (v606:214-215)
 *   mask [N][N] float32 or NULL;  fsc [n][n/2+1] float32 or NULL (ones)
 *   bank [M][n][n/2+1] complex64 out;  norms [M] float32 out or NULL
 *   normalize 1: bank rows unit norm (the reference's bank); 0: rows keep
 *             their scale and norms (required) carries it, for callers that
 *             fold 1/norm into a later step (gis_correlate tinv, the
 *             coefficient gather's scale) and skip a pass over the bank
 */
int gis_build_templates(const void* dft, int N, int n, const float* rts, int M, const float* mask,
                        const float* fsc, void* bank, float* norms, int normalize, void* ws, size_t ws_bytes,
                        void* stream);

/*
 * (c -> e boundary) The template bank reduced to the small-shift search's
 * operand: the transform of gis_build_templates (same arguments and
 * workspace), but direction d stores only the listed packed coefficients,
 *   coeffs[d * Kpad + kslot[r * (n/2+1) + c]] = bank[d][r][c] / norm_d
 * for the entries with kslot >= 0 (slots [0, K)), zeros in [K, Kpad), and
 * norms[d] (or NULL) -- the bank itself is never written.  Equals
 * gis_gather_coefficients(bank with normalize 0, scale 1/norms) bit for bit.
 * Replaces the template side of correlate_templates_and_particle's inputs
 * (v606:1131-1146 bank rows, restricted to the kept coefficients).
 *   kslot [n][n/2+1] int32: slot of each packed entry, -1 if not kept
 *   coeffs [M][Kpad] complex64 out
 */
int gis_build_template_coefficients(const void* dft, int N, int n, const float* rts, int M, const float* mask,
                                    const float* fsc, const int* kslot, int K, int Kpad, void* coeffs, float* norms,
                                    void* ws, size_t ws_bytes, void* stream);

/*
 * Both builders over C candidates' spectra in one launch (the batched
 * engine's template stage, ZL gisapr_batch_search.py:286-291 building one
 * bank per candidate model): spectrum c at dft + c * dft_stride (complex64
 * values), the same rotations, mask and FSC; output row c * M + d (bank,
 * coeffs and norms) for direction d of candidate c.  Results equal C single
 * calls bit for bit.  ws: gis_build_templates_workspace(N, C * M).
 */
int gis_build_templates_batched(const void* dft, size_t dft_stride, int C, int N, int n, const float* rts, int M,
                                const float* mask, const float* fsc, void* bank, float* norms, int normalize, void* ws,
                                size_t ws_bytes, void* stream);

/*
 * gis_build_templates_batched with bank rows `ldb` >= n/2+1 complex values
 * apart ([C*M][n][ldb]); columns past n/2 are written as zeros.  For the
 * searches' aligned operand loads (ldb a multiple of 16).  ldb = 0: the
 * planes in fragment order ([C*M][gis_fragment_plane(n)], see
 * gis_pack_fragments), the searches' 16-byte-load layout.
 */
int gis_build_templates_batched_ld(const void* dft, size_t dft_stride, int C, int N, int n, const float* rts, int M,
                                   const float* mask, const float* fsc, void* bank, int ldb, float* norms,
                                   int normalize, void* ws, size_t ws_bytes, void* stream);
int gis_build_template_coefficients_batched(const void* dft, size_t dft_stride, int C, int N, int n, const float* rts,
                                            int M, const float* mask, const float* fsc, const int* kslot, int K,
                                            int Kpad, void* coeffs, float* norms, void* ws, size_t ws_bytes,
                                            void* stream);

/*
 * Standalone projector (project3d_and_whiten_cuda.py): real-space central
 * slice projections.  Replaces the per-angle loop of its main (:119-147:
 * extract_central_slices_rfft, ifftshift, irfftn, ifftshift, crop) and, with
 * --normal_background_powerspectrum, getSpectrum_divideBySpectrum_torch
 * (:196-230, whiten) then normalize_torch (:231-291, normz: least-squares
 * plane, then (x - mean) / std when std > eps, else x - mean).
 *   dft  [L][L][L/2+1] bricked, from gis_volume_dft_ex(volume, N, pad, ..)
 *   rts  [M][3][3] float32 R = Euler(rot, tilt, 0)^T
 *   proj [M][N][N] float32 out;  ws: gis_build_templates_workspace_ex(N, M, pad)
 */
int gis_project_slices(const void* dft, int N, int pad, const float* rts, int M, int whiten, int normz, float eps,
                       float* proj, void* ws, size_t ws_bytes, void* stream);

/*
 * getSpectrum_divideBySpectrum_torch (whiten) and / or normalize_torch
 * (normz) of project3d_and_whiten_cuda.py:196-291 on M real [N][N] images
 * (at least one of the two).  img, out [M][N][N] float32 (out may not alias
 * img); ws: gis_build_templates_workspace(N, M).
 */
int gis_whiten_images(const float* img, int N, int M, int whiten, int normz, float eps, float* out, void* ws,
                      size_t ws_bytes, void* stream);

/*
 * extract_central_slices_rfft (project3d_and_whiten_cuda.py:397-435; the
 * sampling of v606:1099-1113): trilinear, border-clamped samples of a
 * bricked [L][L][L/2+1] half spectrum at M rotated central slices,
 * conjugated where the rotated x < 0.  rts [M][3][3] act on (x, y, z)
 * frequency vectors; slices [M][L][L/2+1] complex64 out with rows in
 * fftshifted order; L a supported box or twice one;
 * ws: gis_extract_slices_workspace(L, M).
 */
size_t gis_extract_slices_workspace(int L, int M);
int gis_extract_slices(const void* dft, int L, const float* rts, int M, void* slices, void* ws, size_t ws_bytes,
                       void* stream);

/*
 * apply_sinc2_filter_inplace_3d (v606:837-853;
 * project3d_and_whiten_cuda_nowrite.py:42-81): vol[z][y][x] *= sinc^2(|f|)
 * in place, f = fftshift(fftfreq(n)) per axis, vol [D][H][W] float32.
 */
int gis_sinc2_filter(float* vol, int D, int H, int W, void* stream);

/*
 * rotated_grid_to_grid_sample (v606:866-897): for M rotations R [M][3][3]
 * (acting on xyz vectors) of P base frequency points base_zyx [P][3] (z, y, x),
 * the align_corners=True grid_sample coordinates grid_xyz [M][P][3] (x, y, z
 * order) of an rfft volume of shape (dft_d, dft_h, dft_w) sampled as a real
 * image of shape (img_d, img_h, img_w), and conj [M][P] (1 where the rotated
 * x frequency is negative: the point was mirrored and its sample must be
 * conjugated).
 */
int gis_slice_sample_grid(const float* base_zyx, int P, const float* R, int M, int img_d, int img_h, int img_w,
                          int dft_d, int dft_h, int dft_w, float* grid_xyz, unsigned char* conj, void* stream);

/*
 * sample_dft_3d (project3d_and_whiten_cuda.py:612-650): trilinear samples of a
 * complex64 volume dft [D][H][W] (natural layout) at P array coordinates
 * coords [P][3] (z, y, x), each clamped to [0, size-1] (grid_sample with
 * align_corners=True and border padding), into out [P] complex64.
 */
int gis_sample_dft_3d(const void* dft, int D, int H, int W, const float* coords, size_t P, void* out, void* stream);

/* Workspace bytes gis_prepare_particles needs for `particles` (= B) particles
 * of up to 64 psi images each; _ex for P psi images.  The call works through
 * the particles in chunks of the images the workspace holds (at least one
 * particle's P images). */
size_t gis_prepare_particles_workspace(int N, int particles);
size_t gis_prepare_particles_workspace_ex(int N, int particles, int P);

/*
 * (b') Particle preparation.  Replaces ParticleFourierPreparer.prepare
 * (v606:1340-1374) with PackedCTFTorch.batch (v606:778-829) fused in,
 * batched over B particles x P psi angles.
 *   images  [B][N][N] float32
 *   rotp    [B][P][6] float32 psi sampling grids (c, -s, s, c, off_y, off_x),
 *           source = M (y, x) + off, as PsiGridBank v606:1213-1239
 *   mask    [N][N] float32 or NULL
 *   ctf_tab with defocus != NULL: [4][n][n/2+1] float32 (x, y, u4, ncurve of
 *           PackedCTFTorch.__init__), the CTF is evaluated per particle;
 *           with defocus == NULL: [B][n][n/2+1] float32 explicit packed CTF
 *           (e.g. PackedCTFNumPy output), or NULL for no CTF (ones)
 *   defocus [B][3] float32 (dfU, dfV, angle deg) or NULL
 *   kconst  HOST [5] float32: k1, k2, k3, k5, kk
 *   spectra [B][P][n][n/2+1] complex64 out (x CTF, conjugated, unnormalised)
 *   inv_norm [B][P] float32 out: 1 / ||spectrum||_2
 */
int gis_prepare_particles(const float* images, int B, int N, int n, int P, const float* rotp,
                          const float* mask, const float* ctf_tab, const float* defocus,
                          const float* kconst, void* spectra, float* inv_norm, void* ws,
                          size_t ws_bytes, void* stream);

/*
 * gis_prepare_particles with the spectra written rows `ld` >= n/2+1 complex
 * values apart ([B][P][n][ld], columns past n/2 untouched), or with ld = 0 in
 * fragment order ([B][P][gis_fragment_plane(n)], pads zero): the searches'
 * operand layout, written directly (no permuting copy).
 */
int gis_prepare_particles_ld(const float* images, int B, int N, int n, int P, const float* rotp, const float* mask,
                             const float* ctf_tab, const float* defocus, const float* kconst, void* spectra, int ld,
                             float* inv_norm, void* ws, size_t ws_bytes, void* stream);

/*
 * Packed CTF of B particles.  Replaces PackedCTFTorch.batch (v606:778-829).
 *   ctf_tab [4][n][n/2+1] float32 (x, y, u4, ncurve);  defocus [B][3]
 *   kconst  HOST [5] float32 (k1, k2, k3, k5, kk);     out [B][n][n/2+1] float32
 */
int gis_ctf_packed(const float* ctf_tab, const float* defocus, int B, int n, const float* kconst, float* out,
                   void* stream);

/* Workspace bytes gis_correlate needs at newbox n, crop C, B particles. */
size_t gis_correlate_workspace(int n, int C, int B);

/*
 * (d) FFT cross-correlation search + per-particle reduction.  Replaces
 * correlate_templates_and_particle (v606:1418-1449) and the numerics of
 * make_result_records (v606:1452-1491), batched over B particles, on FP32
 * MFMA (the operator as a two-stage linear functional of the unshifted
 * product).  Particle b is correlated against bank rows
 * dir_idx[dir_off[b] .. dir_off[b+1]).
 *   bank [rows][n][ldt], spectra [particles][P][n][ldp] complex64, rows ld
 *        >= n/2+1 complex values apart (columns past n/2 zero when padded);
 *        ld = 0: the planes are in fragment order (gis_pack_fragments),
 *        gis_fragment_plane(n) complex values each
 *   inv_norm [particles][P] 1/norm of each spectrum or NULL; tinv [bank rows]
 *        1/norm of each template row or NULL (bank already unit norm)
 *   pmap [B] spectra row of searched particle b, or NULL (identity): one
 *        launch can search several candidates' banks (rows c*M + direction)
 *        against one resident particle stack (ZL:gisapr_batch_search.py:318-340)
 *   order [B] processing order of the searched particles or NULL
 *   nrows dir_off[B], the number of dir_idx entries
 *   dt, pt directions (1 or 2) / psi values (1..5) per wave, 0 = automatic
 *        (autotuner knobs: results do not depend on them)
 *   scores/peak_y/peak_x [dir_off[B]][P]: score = max(ccg crop) * 1024 n^2,
 *                        (y, x) = divmod(argmax, C) in the C x C crop
 *   best [B]: flat argmax over (local direction, psi) or -1 if none
 *   peak [B] float32, mean/std [B] float64 (population std)
 * C (--ccgCropSize) must be even and <= 32 (the reference default is 32,
 * v606:1562); n <= 256.
 */
int gis_correlate(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                  const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B, int P, int n,
                  int C, int nrows, int ldt, int ldp, int dt, int pt, float* scores, int* peak_y, int* peak_x,
                  int* best, float* peak, double* mean, double* stdv, void* ws, size_t ws_bytes, void* stream);

/*
 * Fragment order: the searches' operand layout, in which each step's 16 x 16
 * coefficient tile is two contiguous 1 KiB runs read with 16-byte loads.
 * A plane of n rows and m = n/2+1 columns is cut into tiles of 16 rows
 * (row steps padded to an even count) by 16 columns, tile by tile; inside a
 * tile, the value at row 16 ks + 4g + 2h + e, column 16 ct + i is at
 * ((h * 4 + g) * 16 + i) * 2 + e.  When m % 16 == 1 the last column is kept
 * apart after the tiles, one value per row.  Rows >= n and columns >= m are
 * zero.  No counterpart in the reference (its planes are row major).
 * gis_fragment_plane: complex values per plane (0 for an odd or tiny n).
 * gis_pack_fragments: dst [rows][gis_fragment_plane(n)] complex64 from src
 * [rows][n][ld] (ld >= n/2+1) on the stream.
 */
size_t gis_fragment_plane(int n);
int gis_pack_fragments(const void* src, int rows, int n, int ld, void* dst, void* stream);

/* Workspace bytes gis_correlate_split needs at newbox n, crop C, B particles. */
size_t gis_correlate_split_workspace(int n, int C, int B);

/*
 * (d) The search of gis_correlate -- the full v606 operator: every packed
 * coefficient, every lag of the C x C crop -- on split-BF16 MFMA (operands
 * split into bf16 hi + lo, hi*hi + hi*lo + lo*hi, FP32 accumulation; ~16
 * significant bits per product).  Replaces correlate_templates_and_particle
 * (v606:1418-1449) and the numerics of make_result_records (v606:1452-1491)
 * like gis_correlate, with the same arguments and outputs, plus
 *   order  [B] processing order of the searched particles (a permutation) or
 *          NULL: particles ranked by orientation share template rows in cache
 *   nrows  dir_off[B], the number of dir_idx entries
 *   ldt    complex values between the rows of a template (n/2+1 for the
 *          [n][n/2+1] layout, or a padded stride: columns past n/2 zero);
 *          a stride that is a multiple of 16 starts every 16-column segment
 *          on a 128-byte line (see gis_build_templates_batched_ld)
 *   ldp    the same for the rows of a psi spectrum; ldt / ldp = 0: the
 *          bank's / spectra's planes are in fragment order (gis_pack_fragments)
 *   dt     local directions per wave (1, 2 or 4; 0 = automatic; crops above
 *          16 use 1)
 *   pt     psi values per wave (1..5; 0 = automatic).  dt and
 *          pt are autotuner knobs: results do not depend on them
 * C (--ccgCropSize) must be even and <= 32 (the reference default is 32,
 * v606:1562); n <= 256.
 */
int gis_correlate_split(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                        const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B, int P,
                        int n, int C, int nrows, int ldt, int ldp, int dt, int pt, float* scores, int* peak_y,
                        int* peak_x, int* best, float* peak, double* mean, double* stdv, void* ws, size_t ws_bytes,
                        void* stream);

/*
 * gis_correlate_split without the workspace argument, allocating nothing:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The workspace
 * is a 1 MiB array per GPU inside the library (held until the process
 * ends); a call whose gis_correlate_split_workspace(n, C, B) exceeds 1 MiB
 * (about 220000 particles at n = 256, C = 32) is refused with GIS_EINVAL.
 * Calls that overlap on one GPU share that array and race: serialise them
 * (one stream, or the caller's own lock).
 */
int gis_correlate_split_noalloc(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                                const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B,
                                int P, int n, int C, int nrows, int ldt, int ldp, int dt, int pt, float* scores,
                                int* peak_y, int* peak_x, int* best, float* peak, double* mean, double* stdv,
                                void* stream);

/*
 * gis_pack_fragments_f16: the psi spectra as the fp16-operand search reads
 * them.  dst [rows][gis_fragment_plane(n)] pairs of fp16 (real, imaginary):
 * the fragment-order planes src [rows][gis_fragment_plane(n)] complex64 times
 * scale * inv_norm[row] (inv_norm may be NULL), rearranged so that a tile's
 * value (row 16 ks + 4g + j, column 16 ct + i) sits at tile * 256 +
 * 4 (16 g + i) + j (one 16-byte run per lane); the lone last column as in
 * fragment order.  No counterpart in the reference (a storage format).
 */
int gis_pack_fragments_f16(const void* src, const float* inv_norm, int rows, int n, float scale, void* dst,
                           void* stream);

/* Workspace bytes gis_correlate_split_f16 needs at newbox n, crop C, B
 * particles and npairs = dir_off[B] * P searched pairs. */
size_t gis_correlate_split_f16_workspace(int n, int C, int B, long npairs);

/*
 * (d) The search of gis_correlate_split with fp16 operands on
 * v_mfma_f32_16x16x32_f16: the same operator (every packed coefficient,
 * every lag of the C x C crop) and outputs, replacing
 * correlate_templates_and_particle (v606:1418-1449) and the numerics of
 * make_result_records (v606:1452-1491).
 *   bank      [M][gis_fragment_plane(n)] complex64, fragment order
 *   bank16    NULL, or bank packed by gis_pack_fragments_f16 with scale 128
 *             and tinv (unit norm x 2^7): the fp16 pass then reads it
 *             instead of rounding the float32 values per step
 *   spectra16 [rows][P][gis_fragment_plane(n)] from gis_pack_fragments_f16
 *             with scale 128 and the spectra's 1/norms (unit norm x 2^7)
 *   spectra32 the float32 fragment-order spectra the fp16 ones were made
 *             from, with their inv_norm, or NULL (see below)
 *   tinv      [M] 1/norm per bank row, or NULL for a unit-norm bank
 *   wsplit    1: the row table W split into fp16 hi + lo (two MFMAs per
 *             tile), 0: fp16 W alone
 * Each step rounds a lane's template values (x 2^7 / norm) to fp16 and forms
 * Q' = T' P' with packed fp16 multiply-adds (Q' = 2^14 Q / (|T| |P|), the A
 * operand as it is); FP32 accumulation; the column stage on split-bf16 MFMA
 * as in gis_correlate_split.  fp16 products keep ~11 significant bits (scores
 * within ~2e-4 of the peak).  With spectra32, each particle's best pair and
 * every pair within max(2^-9 |best|, 64 / sqrt(n (n/2+1))) of it are
 * rescored with the split-bf16 operator of gis_correlate_split before the
 * per-particle tail, so best, peak and the best pair's lag are those of the
 * split-bf16 search.  dir_off, dir_idx, pmap, order, nrows, dt, pt and the
 * outputs as in gis_correlate_split; ws: gis_correlate_split_f16_workspace
 * bytes; C even and <= 32, n <= 256.
 */
int gis_correlate_split_f16(const void* bank, const void* bank16, const void* spectra16, const void* spectra32,
                            const float* inv_norm, const float* tinv, const int* dir_off, const int* dir_idx,
                            const int* pmap, const int* order, int B, int P, int n, int C, int nrows, int dt, int pt,
                            int wsplit, float* scores, int* peak_y, int* peak_x, int* best, float* peak,
                            double* mean, double* stdv, void* ws, size_t ws_bytes, void* stream);

/*
 * Cross-particle conformation score sum_b exp(peak_b / 1024) in float64 over
 * particles with best_b >= 0 (best may be NULL).  Replaces
 * new_method_to_fit_the_2nd_Gaussian_PEAK_v4.py:133-134.  out: 1 double.
 */
int gis_score_sum(const float* peak, const int* best, int B, double* out, void* stream);

/*
 * Per-candidate conformation scores in one launch (ZL:gisapr_batch_search.py:361):
 * out[s] = sum over i in [off[s], off[s+1]) with best_i >= 0 of exp(peak_i / 1024),
 * float64.  off [nseg+1] device int32; out [nseg].
 */
int gis_score_sum_segments(const float* peak, const int* best, const int* off, int nseg, double* out,
                           void* stream);

/*
 * Gathers K listed coefficients (flat packed indices kidx) of `rows` spectra
 * of `plane` complex64 values each into dst [rows][Kpad] complex64, zero
 * padded, times scale[row] if scale is not NULL.  Feeds the small-shift
 * search below; it takes the place of the full-plane products of
 * correlate_templates_and_particle (v606:1418-1449) for coefficients outside
 * the resolution disk.
 */
int gis_gather_coefficients(const void* src, size_t plane, int rows, const int* kidx, int K, int Kpad,
                            const float* scale, void* dst, void* stream);

/*
 * Small-shift precision variant of correlate_templates_and_particle
 * (v606:1418-1449) and the numerics of make_result_records (v606:1452-1491):
 * the score at lag (ly, lx), |ly|, |lx| <= half <= 3, is evaluated as the
 * linear functional 1024 * sum_k Re(T_k P_k Phi[k][lag]) over the gathered
 * coefficients on split-bf16 MFMA (hi*hi + hi*lo + lo*hi, FP32 accumulation).
 *   Tk [bank rows][Kpad], Pk [particles*P][Kpad] complex64 (unit norm over the full plane)
 *   pmap [B] Pk particle row of searched particle b, or NULL (identity)
 *   rows [nrows] processing order of the dir_idx rows (a permutation) or NULL;
 *        grouping one particle's rows of several candidates lets them share
 *        its Pk rows in cache
 *   nrows = dir_off[B] (the number of dir_idx entries)
 *   Bhi / Blo [4][Kpad/16][64][8] bf16: split lag table in MFMA B-operand order
 *   Plast [Kpad] float2 (Re Phi, -Im Phi) of the last lag: when the lag count
 *         is 16 t + 1 (49 for half = 3) t tiles run on MFMA and the last lag
 *         as a float32 dot product; NULL otherwise
 *   pt  psi values per wave (1..5, 0 = min(P, 5)); wpb waves per block (2 or 4; 0 = 4)
 *   scores/peak_y/peak_x [dir_off[B]][P] (peak indices are lag + half),
 *   best/peak/mean/std [B] as gis_correlate (FP64 mean/std).
 */
int gis_correlate_mixed(const void* Tk, const void* Pk, const int* dir_off, const int* dir_idx, const int* pmap,
                        const int* rows, int B, int P, int nrows, int Kpad, const void* Bhi, const void* Blo, const void* Plast,
                        int half, int pt, int wpb, float* scores, int* peak_y, int* peak_x, int* best, float* peak,
                        double* mean, double* stdv, void* stream);

/* 16-bit operand type of gis_correlate_mixed_ex. */
enum { GIS_OPERAND_BF16 = 0, GIS_OPERAND_F16 = 1 };

/*
 * gis_correlate_mixed with a choice of operand type (the precision variant's
 * "FP16 variant"; same reference functional as gis_correlate_mixed).
 * operand = GIS_OPERAND_BF16: as gis_correlate_mixed.  GIS_OPERAND_F16: Bhi /
 * Blo hold the lag table split into fp16 hi + lo (same layout), the A operand
 * is scaled by 2^14 before its fp16 split (|T_k P_k| <= 1, so it stays finite)
 * and the accumulator by 2^-14 (v_mfma_f32_16x16x32_f16, ~22 significant
 * bits per product against split-bf16's ~16).
 */
int gis_correlate_mixed_ex(const void* Tk, const void* Pk, const int* dir_off, const int* dir_idx, const int* pmap,
                           const int* rows, int B, int P, int nrows, int Kpad, const void* Bhi, const void* Blo,
                           const void* Plast, int half, int pt, int wpb, float* scores, int* peak_y, int* peak_x,
                           int* best, float* peak, double* mean, double* stdv, int operand, void* stream);

/*
 * Geometric restraint (func_check_boundary_for_testing_v8.py:62-126):
 * gis_threshold_mask: mask[i] = vol[i] > thr (the body / chain masks, :84, :112).
 * gis_edt_squared: squared Euclidean distance (voxels^2) from every voxel of
 *   an [N]^3 grid to the nearest voxel with mask != 0, exact (1e30 when the
 *   mask is empty); replaces scipy.ndimage.distance_transform_edt(~mask)
 *   (:90-92) up to the square root.  N <= 1024.
 * gis_restraint_eval: one pose of GeometryRestraint.evaluate (:96-126):
 *   overlap[0] = #(chain_vol > thr and main_mask), min_d2[0] = float32 bits of
 *   the smallest main_d2 over voxels with chain_vol > thr (1e30 if none),
 *   any[0] = 1 if some voxel has chain_vol > thr.  Device outputs, set here.
 */
int gis_threshold_mask(const float* vol, size_t n, float thr, unsigned char* mask, void* stream);
int gis_edt_squared(const unsigned char* mask, int N, float* d2, void* stream);
int gis_restraint_eval(const float* chain_vol, float thr, const unsigned char* main_mask, const float* main_d2,
                       size_t n, unsigned long long* overlap, unsigned int* min_d2, int* any, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GISAPR_HIP_H */

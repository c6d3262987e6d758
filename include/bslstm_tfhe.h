/*
 * bslstm_tfhe.h -- C ABI of the MI355X (gfx950) TFHE bootstrapping engine.
 *
 * Plain C types only: device pointers are passed as typed pointers, the HIP
 * stream as `void *` (a hipStream_t; NULL = the default stream).  Every entry
 * point is asynchronous on the given stream, allocates no device memory,
 * validates its shape/attribute arguments on the host and returns a status
 * code; bslstm_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Torus values are integers mod q = 2^log2_q.  Complex128 arrays are
 * interleaved (re, im) doubles.
 *
 * Each entry point names the reference interface it replaces
 * (AMIAD-Research/Blind-Spiking-LSTM, file:line).
 */
#ifndef BSLSTM_TFHE_H
#define BSLSTM_TFHE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Mirrors the XLA FFI error kinds the reference handler
 * returns: ffi::Error::InvalidArgument (csrc/mul_bootstrap.cuh:212,223) and
 * ffi::Error::Internal on a failed launch (csrc/mul_bootstrap.cuh:183-188). */
enum {
    BSLSTM_OK = 0,
    BSLSTM_INVALID_ARGUMENT = 1,
    BSLSTM_INTERNAL = 2
};

/* Message of the last failed call on this thread ("" if none). */
const char *bslstm_last_error(void);

/* ABI version (incremented on any signature change). */
int bslstm_abi_version(void);

/*
 * Blind rotation, plain single-launch FP64 form: one workgroup per
 * ciphertext runs the whole collapsed key-unrolled loop.
 *
 * Replaces the XLA FFI handler `mul_bootstrap` (csrc/mul_bootstrap.cuh:192-243,
 * launcher :138-190, kernel :97-136) registered by cuTFHE/multiply.py:10-14.
 * Operands in the handler's order and layout:
 *   ct_init     int64 (..., 2, N); ct_elems = its element count.  If
 *               ct_elems > 2N every ciphertext has its own test polynomial
 *               (stride 2N), else one is broadcast (:162-164).
 *   bsk         complex128 (n_iter, a=2, l, c=2, m, N/2); bsk_dims = the six
 *               trailing dimensions (:149-156).
 *   all_rot_fft complex128 (2N, N/2) FFT of X^r, or NULL.  Its content is a
 *               fixed function of N (parameters_crypto.py:86-91); the kernel
 *               evaluates the same roots exactly and only checks the shape.
 *   indices     int32 (batch, n_iter, m); idx_elems = its element count.
 *   output      int64 (batch, 2, N).
 * Attributes log2_beta, log2_q, l, degree as in the handler (:238-241).
 * degree must be a power of two in [128, 4096] and l in [1, 4]
 * ("Unsupported degree" / "Unsupported l", :211-224).
 */
int bslstm_mul_bootstrap(void *stream,
                         const int64_t *ct_init, int64_t ct_elems,
                         const double *bsk, const int64_t bsk_dims[6],
                         const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                         const int32_t *indices, int64_t idx_elems,
                         int64_t *output, int64_t batch,
                         int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree);

/*
 * Blind rotation, batched FP64 form: same operands, attributes, semantics and
 * errors as bslstm_mul_bootstrap, and bit-identical outputs; each workgroup
 * runs two ciphertexts and feeds every bootstrapping-key value it loads into
 * both, halving the key traffic from L2.  This is the engine's form of the
 * reference handler's work (csrc/mul_bootstrap.cuh:97-243); degree 4096 runs
 * one ciphertext per workgroup.
 */
int bslstm_mul_bootstrap_batched(void *stream,
                                 const int64_t *ct_init, int64_t ct_elems,
                                 const double *bsk, const int64_t bsk_dims[6],
                                 const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                                 const int32_t *indices, int64_t idx_elems,
                                 int64_t *output, int64_t batch,
                                 int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree);

/*
 * Same contract, operands, errors and bit-identical outputs as
 * bslstm_mul_bootstrap (csrc/mul_bootstrap.cuh:138-243); the latency form:
 * one ciphertext per workgroup, its two accumulator rows (mask, body) split
 * over two thread groups of N/8 threads, so each step's transforms and
 * multiply-accumulate run on twice the threads.  For small batches, where the
 * launch is shorter than one wave of workgroups.  Degrees 128, 256 and 4096
 * run the one-group form.
 */
int bslstm_mul_bootstrap_rows(void *stream,
                              const int64_t *ct_init, int64_t ct_elems,
                              const double *bsk, const int64_t bsk_dims[6],
                              const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                              const int32_t *indices, int64_t idx_elems,
                              int64_t *output, int64_t batch,
                              int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree);

/*
 * The "pair" latency form: same contract, operands, errors and bit-identical
 * outputs as bslstm_mul_bootstrap (csrc/mul_bootstrap.cuh:138-243), over two
 * workgroups (two CUs) per ciphertext.  Each workgroup keeps one accumulator
 * row, transforms that row's gadget digits, exchanges the spectra with its
 * partner through device memory once per CMux step and multiplies half of
 * the step's key slice: for a few dozen ciphertexts, where one CU per
 * ciphertext is bound by reading the key slice from L2.
 *
 * Degree 1024 or 2048 and l <= 3; at most CU count / 2 ciphertexts at
 * N = 2048 (one workgroup per CU) and CU count at N = 1024 (two per CU),
 * rounded up to a multiple of 8, so that every workgroup is resident at once.  Two
 * pair-form launches that run concurrently (on different streams) share that
 * budget: their ciphertexts together should not exceed it, or waits may run
 * out (reported in the status word; never a hang).
 * workspace: device memory of bslstm_mul_bootstrap_pair_workspace_bytes()
 * bytes, cleared by the call itself (stream-ordered; one call at a time per
 * workspace).  Its first 4 bytes are a status word the caller may read after
 * the call: bit 0 = a wait on the partner workgroup ran out after
 * spin_limit polls (the outputs are then wrong; the kernel still ends),
 * bit 1 = some pair ran on two XCDs (slower coherent exchange, same result).
 * spin_limit in [1, 2^30]; each poll sleeps ~50 ns.
 */
int64_t bslstm_mul_bootstrap_pair_workspace_bytes(int64_t batch, int64_t l, int64_t degree);

int bslstm_mul_bootstrap_pair(void *stream,
                              const int64_t *ct_init, int64_t ct_elems,
                              const double *bsk, const int64_t bsk_dims[6],
                              const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                              const int32_t *indices, int64_t idx_elems,
                              int64_t *output, int64_t batch,
                              int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree,
                              void *workspace, int64_t workspace_bytes, int64_t spin_limit);

/* bslstm_mul_bootstrap_pair with bsk_pool (slot_count, n_iter, 2, l, 2, m,
 * N/2) and slots int32 (batch), as bslstm_mul_bootstrap_slots. */
int bslstm_mul_bootstrap_pair_slots(void *stream,
                                    const int64_t *ct_init, int64_t ct_elems,
                                    const double *bsk_pool, const int64_t bsk_dims[6],
                                    const int32_t *slots, int64_t slot_count,
                                    const int32_t *indices, int64_t idx_elems,
                                    int64_t *output, int64_t batch,
                                    int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree,
                                    void *workspace, int64_t workspace_bytes, int64_t spin_limit);

/* ------------------------------------------------------------------------
 * Blind rotation in exact integer arithmetic on the matrix cores (the
 * "split" form): the same computation as bslstm_mul_bootstrap (the handler
 * csrc/mul_bootstrap.cuh:192-243, kernel :20-136), with every negacyclic
 * product exact -- the bootstrapping key cut into signed byte limbs, the
 * batched products int8 MFMA GEMMs with int32 accumulation -- so outputs
 * equal the exact (int128) product mod q bit for bit.  Needs log2_q <= 32,
 * log2_beta <= 7 and degree 512, 1024 or 2048.
 * ------------------------------------------------------------------------ */

/* Bytes of the prepared key for bsk_dims (n_iter, 2, l, 2, m, N/2) at this
 * degree (-1 on bad arguments). */
int64_t bslstm_mul_bootstrap_split_key_bytes(const int64_t bsk_dims[6], int64_t degree);

/* The FFT-domain key of bslstm_mul_bootstrap (complex128 (n_iter, 2, l, 2,
 * m, N/2), multiply.py:81-86 layout; its integer coefficients are recovered
 * exactly) -> split_key of bslstm_mul_bootstrap_split_key_bytes() bytes.
 * Once per key. */
int bslstm_mul_bootstrap_split_prepare_key(void *stream, const double *bsk,
                                           const int64_t bsk_dims[6], int64_t l, int64_t degree,
                                           int8_t *split_key);

/* Workspace for a batch processed in chunks of up to `chunk` ciphertexts
 * (chunk <= 0: 512); a larger workspace only allows larger chunks. */
int64_t bslstm_mul_bootstrap_split_workspace_bytes(int64_t batch, const int64_t bsk_dims[6],
                                                   int64_t degree, int64_t chunk);

/* bslstm_mul_bootstrap's operands, attributes and errors with split_key in
 * place of the FFT-domain key (all_rot_fft is not needed) and a workspace. */
int bslstm_mul_bootstrap_split(void *stream, const int64_t *ct_init, int64_t ct_elems,
                               const int8_t *split_key, const int64_t bsk_dims[6],
                               const int32_t *indices, int64_t idx_elems, int64_t *output,
                               int64_t batch, int64_t log2_beta, int64_t log2_q, int64_t l,
                               int64_t degree, void *workspace, int64_t workspace_bytes);

/*
 * Rotation exponents of the collapsed blind rotation.
 * Replaces the index preparation of cuTFHE/multiply.py:67-79 (an XLA fusion
 * in the reference):
 *   val[g,p] = sum_j pk[g*c+j] * bit_j(p)        (bit order of cuTFHE/utils.py:5-8)
 *   idx      = int32(rint(2N*val/(q*n_lut)) * n_lut) mod 2N   (rint half-even,
 *              floor modulo)
 * public_key float64 (batch, n); indices int32 (batch, n/collapse, 2^collapse).
 */
int bslstm_rotation_indices(void *stream, const double *public_key, int64_t batch, int64_t n,
                            int64_t collapse, int64_t log2_q, int64_t degree, int64_t n_lut,
                            int32_t *indices);

/*
 * Many-LUT tail of the programmable bootstrap: rotate each blind-rotated
 * test polynomial by X^(-b_ms), b_ms = rint(2N*b/(q*n_lut))*n_lut, and
 * sample-extract coefficients 0..n_lut-1.
 * Replaces cuboot_merge's tail, schemas/RLWE_jax.py:295-303, with
 * rotate_ciphertext (schemas/text_jax.py:348-362, coef_rotation
 * schemas/polynomial_jax.py:150-169) and sample_extract
 * (schemas/text_jax.py:191-208).
 *   br_out  int64 (batch, 2, N);  body float64 (batch)  (the LWE bodies b)
 *   masks   float64 (batch, n_lut, N);  bodies float64 (batch, n_lut)
 */
int bslstm_rotate_extract(void *stream, const int64_t *br_out, const double *body,
                          int64_t batch, int64_t degree, int64_t log2_q, int64_t n_lut,
                          double *masks, double *bodies);

/*
 * LWE key switch.
 * Replaces key_switch_LWE + gadget_product_LWE (schemas/LWE_jax.py:97-117,
 * 65-82) with the float digit rule of decomposition
 * (schemas/polynomial_jax.py:90-111):
 *   d[j,i] = digits of a[i] (j < l_ks, round half-even, base 2^log2_beta)
 *   a'     = centred_mod(sum_{j,i} d[j,i] * ksk_a[j,i,:], q)
 *   b'     = centred_mod(b + centred_mod(sum_{j,i} d[j,i] * ksk_b[j,i], q), q)
 * a_in float64 (batch, n_in), b_in float64 (batch); ksk_a float64
 * (l_ks, n_in, n_out), ksk_b float64 (l_ks, n_in) as built by
 * get_keyswitch_sk_LWE (schemas/LWE_jax.py:86-93); a_out float64
 * (batch, n_out), b_out float64 (batch).  Requires log2_q <= 32 so that the
 * sums are exact in FP64 (the reference's own tensordot is exact there).
 */
int bslstm_key_switch_lwe(void *stream, const double *a_in, const double *b_in,
                          const double *ksk_a, const double *ksk_b,
                          int64_t batch, int64_t n_in, int64_t n_out,
                          int64_t log2_beta_ks, int64_t l_ks, int64_t log2_q,
                          double *a_out, double *b_out);

/* ------------------------------------------------------------------------
 * int64 forms.  The reference runs its LSTM at q = 2^64 (parameters_crypto.py:6)
 * with float64 ciphertexts, which cannot hold 64-bit classes; these entry
 * points take int64 centred representatives (any class representative is
 * accepted on input) for any log2_q <= 64, with integer sums wrapping mod 2^64.
 * ------------------------------------------------------------------------ */

/* bslstm_rotation_indices with int64 masks: the group sums are exact, then
 * the same FP64 rounding (cuTFHE/multiply.py:67-79). */
int bslstm_rotation_indices_i64(void *stream, const int64_t *public_key, int64_t batch, int64_t n,
                                int64_t collapse, int64_t log2_q, int64_t degree, int64_t n_lut,
                                int32_t *indices);

/* bslstm_rotate_extract with int64 operands and a separate extraction count:
 * rotate by X^(-b_ms), b_ms = rint(2N*b/(q*n_lut))*n_lut (no rotation when
 * body is NULL), then sample-extract coefficients 0..n_ext-1.
 * Replaces schemas/RLWE_jax.py:295-303 (n_ext = n_lut) and the hidden-state
 * extraction vmap(sample_extract)(C, arange(hidden)) of
 * nn_tfhe/Spike_LSTM.py:113-114 (body NULL, n_ext = hidden).
 *   br_out int64 (batch, 2, N); body int64 (batch) or NULL;
 *   masks int64 (batch, n_ext, N); bodies int64 (batch, n_ext). */
int bslstm_rotate_extract_i64(void *stream, const int64_t *br_out, const int64_t *body,
                              int64_t batch, int64_t degree, int64_t log2_q, int64_t n_lut,
                              int64_t n_ext, int64_t *masks, int64_t *bodies);

/* Exact LWE key switch for log2_q <= 64 (schemas/LWE_jax.py:97-117, 65-82;
 * digits by the float rule of schemas/polynomial_jax.py:90-111 restated on
 * integers).  Layouts as bslstm_key_switch_lwe, all int64. */
int bslstm_key_switch_lwe_i64(void *stream, const int64_t *a_in, const int64_t *b_in,
                              const int64_t *ksk_a, const int64_t *ksk_b,
                              int64_t batch, int64_t n_in, int64_t n_out,
                              int64_t log2_beta_ks, int64_t l_ks, int64_t log2_q,
                              int64_t *a_out, int64_t *b_out);

/*
 * Homomorphic Linear layer with coefficient-0 extraction.
 * Replaces Linear.__call__ (nn_tfhe/Spike_LSTM.py:46-61): for every output
 * row o, multiply_plaintext_ciphertext(W_o, ct) (schemas/text_jax.py:59-78),
 * sample_extract(., 0) (text_jax.py:191-208), + bias_o, centred mod q.
 * The product is exact (the reference rounds an FP64 FFT product).
 *   ct      int64 (batch, 2, N) or (2, N) broadcast (ct_elems = its size)
 *   weights int64 (out_dim, N) weight polynomials (Linear.set_weights,
 *           Spike_LSTM.py:27-41); support int32 (K) the columns that may be
 *           nonzero in any row (all others are skipped; entries outside
 *           [0, N) are ignored on the device); bias int64 (out_dim) or NULL
 *   masks   int64 (batch, out_dim, N); bodies int64 (batch, out_dim)
 */
int bslstm_linear_extract(void *stream, const int64_t *ct, int64_t ct_elems,
                          const int64_t *weights, const int32_t *support, int64_t K,
                          const int64_t *bias, int64_t batch, int64_t out_dim,
                          int64_t degree, int64_t log2_q, int64_t *masks, int64_t *bodies);

/*
 * LWE -> RLWE trace packing, one RLWE per LWE.
 * Replaces pack_lwe_to_rlwe (schemas/text_jax.py:314-345) with
 * modulus_switch (:159-179), trivial_embedding (:268-287) and
 * apply_automorphism_and_keyswitch (:290-311).
 *   lwe_a int64 (count, N), lwe_b int64 (count)
 *   galois_key_fft complex128 (log2 N, l, 2, N/2): get_galois_key
 *           (schemas/RLWE_jax.py:217-226), reference bin order
 *   out int64 (count, 2, N)
 * Digits use log2_beta / l (dict_params_packing beta_ks / l_ks).
 */
int bslstm_pack_lwe_to_rlwe(void *stream, const int64_t *lwe_a, const int64_t *lwe_b,
                            int64_t count, const double *galois_key_fft, int64_t degree,
                            int64_t log2_beta, int64_t l, int64_t log2_q, int64_t *out);

/*
 * Plaintext-LUT products summed over the many-LUT slots, for a batch of
 * sequences sharing the LUTs.
 * Replaces vmap(vmap(prepare_LUT))'s product (nn_tfhe/activation.py:103-106,
 * multiply_plaintext_ciphertext with fft_product_twice) and the slot sum of
 * nn_tfhe/Spike_LSTM.py:129-132:
 *   out[s, g] = centred(sum_k round(LUT[k, g] * RLWE[s, k, g]))
 *   rlwe int64 (batch, n_lut, G, 2, N); lut_fft complex128 (n_lut, G, N/2)
 *   (the LUT polynomials through bslstm_poly_fft); out int64 (batch, G, 2, N).
 */
int bslstm_lut_product_sum(void *stream, const int64_t *rlwe, const double *lut_fft,
                           int64_t batch, int64_t groups, int64_t n_lut, int64_t degree,
                           int64_t log2_q, int64_t *out);

/* Negacyclic transform of integer polynomials into complex128 (count, N/2),
 * bins F_j = p(zeta^(2j+1)), zeta = exp(-i pi/N): jax_fourier
 * (schemas/polynomial_jax.py:11-15) for keys and LUTs. */
int bslstm_poly_fft(void *stream, const int64_t *polys, int64_t count, int64_t degree,
                    double *out);

/* Inverse of bslstm_poly_fft: spectra complex128 (count, N/2) in the same
 * bin order -> int64 (count, N), each coefficient rounded to the nearest
 * integer (jax_inversefourier, schemas/polynomial_jax.py:17-21). */
int bslstm_poly_ifft(void *stream, const double *spectra, int64_t count, int64_t degree,
                     int64_t *out);

/* Workspace of bslstm_packing in bytes (-1 on bad arguments). */
int64_t bslstm_packing_workspace_bytes(int64_t packs, int64_t n_in, int64_t degree);

/*
 * Packing of up to N LWEs into one RLWE with the packing key switching key.
 * Replaces packing (schemas/text_jax.py:211-239); `packs` independent
 * packings per call.
 *   masks int64 (packs, n_lwes, n_in), bodies int64 (packs, n_lwes),
 *   n_lwes <= N; LWE i lands on coefficient i.
 *   ksk_fft complex128 (n_in, l, 2, N/2): get_packing_KSK
 *           (schemas/RLWE_jax.py:205-214) through bslstm_poly_fft
 *   workspace device memory of bslstm_packing_workspace_bytes() bytes
 *   out int64 (packs, 2, N)
 * The transform-domain sum over the n_in rows is rounded once (the
 * reference rounds each row's product before summing).
 */
int bslstm_packing(void *stream, const int64_t *masks, const int64_t *bodies, int64_t packs,
                   int64_t n_lwes, int64_t n_in, const double *ksk_fft, int64_t degree,
                   int64_t log2_beta, int64_t l, int64_t log2_q, void *workspace,
                   int64_t workspace_bytes, int64_t *out);

/* out = centred(sx*x + sy*y mod q) elementwise over n int64 values:
 * sum_ciphertext_ciphertext (schemas/text_jax.py:27-43) and its differences. */
int bslstm_torus_axpy(void *stream, const int64_t *x, const int64_t *y, int64_t n, int64_t sx,
                      int64_t sy, int64_t log2_q, int64_t *out);

/* out[r] = X^e * in[r] in Z_q[X]/(X^N+1) for rows r < rows (int64 (rows, N)),
 * e = exponents[r / rows_per_exponent]: coef_rotation
 * (schemas/polynomial_jax.py:150-169) / rotate_ciphertext
 * (schemas/text_jax.py:348-362).  in and out must not alias. */
int bslstm_rotate_poly(void *stream, const int64_t *in, int64_t rows, int64_t degree,
                       const int64_t *exponents, int64_t rows_per_exponent, int64_t log2_q,
                       int64_t *out);

/* ------------------------------------------------------------------------
 * LWE key switch on the matrix cores (the exact int8 "split" form).
 * Same mathematics as bslstm_key_switch_lwe_i64 (schemas/LWE_jax.py:97-117)
 * and bit-identical results: the digit matrix (int8) times the key cut into
 * eight balanced signed bytes, as int8 MFMA GEMMs with int32 accumulation,
 * recombined exactly mod 2^64.  Needs log2_beta_ks <= 7 (int8 digits) and
 * n_in*l_ks < 2^18 (exact int32 sums).
 * ------------------------------------------------------------------------ */

/* Bytes of the prepared key (-1 on bad arguments). */
int64_t bslstm_ks_mfma_key_bytes(int64_t l, int64_t n_in, int64_t n_out);

/* Cut ksk_a int64 (l, n_in, n_out) and ksk_b int64 (l, n_in) (as built by
 * get_keyswitch_sk_LWE, schemas/LWE_jax.py:86-93) into the int8 limb tiles
 * the GEMM reads; key_tiles holds bslstm_ks_mfma_key_bytes() bytes.  Once
 * per key. */
int bslstm_ks_mfma_prepare_key(void *stream, const int64_t *ksk_a, const int64_t *ksk_b,
                               int64_t l, int64_t n_in, int64_t n_out, int8_t *key_tiles);

/* Workspace of bslstm_key_switch_lwe_mfma in bytes (the int8 digit tiles). */
int64_t bslstm_ks_mfma_workspace_bytes(int64_t batch, int64_t l, int64_t n_in);

/* a_in int64 (batch, n_in), b_in int64 (batch) -> a_out int64 (batch, n_out),
 * b_out int64 (batch), batch <= 2097120 (65535 row tiles of 32); key_tiles
 * from bslstm_ks_mfma_prepare_key. */
int bslstm_key_switch_lwe_mfma(void *stream, const int64_t *a_in, const int64_t *b_in,
                               const int8_t *key_tiles, int64_t batch, int64_t n_in,
                               int64_t n_out, int64_t log2_beta_ks, int64_t l_ks, int64_t log2_q,
                               void *workspace, int64_t workspace_bytes,
                               int64_t *a_out, int64_t *b_out);

/*
 * Allocation-free form of bslstm_key_switch_lwe_mfma.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same operands, results (bit-identical) and errors as
 * bslstm_key_switch_lwe_mfma minus the workspace: the int8 digit tiles live
 * in a static per-device array of bslstm_key_switch_lwe_mfma_noalloc_max_bytes()
 * bytes (16 MiB), so batches whose bslstm_ks_mfma_workspace_bytes() exceed it
 * are rejected with BSLSTM_INVALID_ARGUMENT.  The call allocates nothing.
 * All calls on one device share that array: calls that overlap on one
 * device (concurrent streams or host threads) race.
 */
int64_t bslstm_key_switch_lwe_mfma_noalloc_max_bytes(void);

int bslstm_key_switch_lwe_mfma_noalloc(void *stream, const int64_t *a_in, const int64_t *b_in,
                                       const int8_t *key_tiles, int64_t batch, int64_t n_in,
                                       int64_t n_out, int64_t log2_beta_ks, int64_t l_ks,
                                       int64_t log2_q, int64_t *a_out, int64_t *b_out);

/* ------------------------------------------------------------------------
 * Key-slot forms: one launch over a batch that mixes clients.  Each client's
 * key sits in one slot of a device pool of equally shaped keys; every item
 * names its slot (int32, clamped to [0, slot_count) on the device).  No
 * counterpart in the reference, which serves one key set per process: these
 * are for the continuous-batching server (requests of many clients share
 * every launch of a timestep).
 * ------------------------------------------------------------------------ */

/* bslstm_mul_bootstrap(_batched, _rows) with bsk_pool (slot_count, n_iter,
 * 2, l, 2, m, N/2) and slots int32 (batch).  form 1 = one ciphertext per
 * workgroup, 2 = two (then ciphertexts 2i and 2i+1 must share a slot),
 * 3 = one ciphertext over two row groups; other values: InvalidArgument. */
int bslstm_mul_bootstrap_slots(void *stream, const int64_t *ct_init, int64_t ct_elems,
                               const double *bsk_pool, const int64_t bsk_dims[6],
                               const int32_t *slots, int64_t slot_count,
                               const int32_t *indices, int64_t idx_elems, int64_t *output,
                               int64_t batch, int64_t log2_beta, int64_t log2_q, int64_t l,
                               int64_t degree, int64_t form);

/* bslstm_key_switch_lwe_mfma with key_pool (slot_count x key bytes) and
 * tile_slots int32 (ceil(batch/32)): rows 32t .. 32t+31 use slot
 * tile_slots[t]. */
int bslstm_key_switch_lwe_mfma_slots(void *stream, const int64_t *a_in, const int64_t *b_in,
                                     const int8_t *key_pool, const int32_t *tile_slots,
                                     int64_t slot_count, int64_t batch, int64_t n_in,
                                     int64_t n_out, int64_t log2_beta_ks, int64_t l_ks,
                                     int64_t log2_q, void *workspace, int64_t workspace_bytes,
                                     int64_t *a_out, int64_t *b_out);

/* bslstm_pack_lwe_to_rlwe with galois_key_pool (slot_count, log2 N, l, 2,
 * N/2) and slots int32 (count). */
int bslstm_pack_lwe_to_rlwe_slots(void *stream, const int64_t *lwe_a, const int64_t *lwe_b,
                                  int64_t count, const double *galois_key_pool,
                                  const int32_t *slots, int64_t slot_count, int64_t degree,
                                  int64_t log2_beta, int64_t l, int64_t log2_q, int64_t *out);

/* bslstm_packing with ksk_pool (slot_count, n_in, l, 2, N/2) and slots int32
 * (packs). */
int bslstm_packing_slots(void *stream, const int64_t *masks, const int64_t *bodies,
                         int64_t packs, int64_t n_lwes, int64_t n_in, const double *ksk_pool,
                         const int32_t *slots, int64_t slot_count, int64_t degree,
                         int64_t log2_beta, int64_t l, int64_t log2_q, void *workspace,
                         int64_t workspace_bytes, int64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* BSLSTM_TFHE_H */

/* fx_trig.hpp -- sine and cosine of a float, written out as a fixed
 * sequence of IEEE operations, so that the kernels and any restatement of
 * them on a CPU compute the same bits.  Used for the moving X-min wall of the
 * advance stage (fluids5.0/src/particles.cu:275: sin(time * freq + z / 400))
 * and its adjoint.  The C library's sinf and the GPU math library's each
 * round correctly most of the time but not always the same way; one ulp of
 * difference in the wall position is enough to make two otherwise identical
 * trajectories part.
 *
 * Method (textbook argument reduction and Taylor series), all in float:
 *   k = rint(x * 2/pi);  r = x - k * pi/2 with pi/2 = P1 + P2 + P3, three floats
 *   (Cody and Waite's split), each product folded in by one fmaf, so that the
 *   first step is exact and r is accurate to about an ulp for |k| up to
 *   millions; |r| <= pi/4 plus rounding;
 *   sin r = r - r^3/3! + r^5/5! - r^7/7! + r^9/9!   (truncation < 1.8e-9)
 *   cos r = 1 - r^2/2! + ... - r^10/10!              (truncation < 1.2e-10)
 *   by Horner's rule with fmaf; k mod 4 (fmodf, exact for any float k)
 *   selects sin or cos of r and the sign.  The constants are the floats
 *   nearest to 2/pi and to the successive remainders of pi/2 (from a 60-digit
 *   pi), and the floats nearest to 1/n!.  No double arithmetic: the device
 *   runs it in FP32.
 * Accuracy against a double-precision sine and cosine: within 2 ulps of the
 * float result for |x| up to 1e6 (checked by tests/test_oracle_trig.py).
 * Infinite and NaN arguments give NaN; sin(-0) gives +0.
 */
#ifndef FX_TRIG_HPP
#define FX_TRIG_HPP

#include <math.h>

#if defined(__HIPCC__)
#define FX_TRIG_FN __host__ __device__ inline
#else
#define FX_TRIG_FN inline
#endif

/* sin(x) for shift 0, cos(x) = sin(x + pi/2) for shift 1 */
FX_TRIG_FN float fx_trig_shifted(float x, int shift) {
  if (!(fabsf(x) <= 0x1.fffffep+127f)) return x - x; /* NaN for +-inf and NaN */
  const float k = rintf(x * 0x1.45f306p-1f);          /* 2/pi */
  float r = fmaf(-k, 0x1.921fb6p+0f, x);               /* pi/2 = P1 */
  r = fmaf(-k, -0x1.777a5cp-25f, r);                   /*       + P2 */
  r = fmaf(-k, -0x1.ee59dap-50f, r);                   /*       + P3 */
  const int q = ((int)fmodf(k, 4.0f) + shift) & 3;
  const float r2 = r * r;
  float s = fmaf(0x1.71de3ap-19f, r2, -0x1.a01a02p-13f); /* 1/9!, -1/7! */
  s = fmaf(s, r2, 0x1.111112p-7f);                        /* 1/5! */
  s = fmaf(s, r2, -0x1.555556p-3f);                       /* -1/3! */
  s = fmaf(r * r2, s, r);
  float c = fmaf(-0x1.27e4fcp-22f, r2, 0x1.a01a02p-16f); /* -1/10!, 1/8! */
  c = fmaf(c, r2, -0x1.6c16c2p-10f);                      /* -1/6! */
  c = fmaf(c, r2, 0x1.555556p-5f);                        /* 1/4! */
  c = fmaf(c, r2, -0.5f);                                 /* -1/2! */
  c = fmaf(r2, c, 1.0f);
  const float v = (q & 1) ? c : s;
  return (q & 2) ? -v : v;
}

FX_TRIG_FN float fx_sinf(float x) { return fx_trig_shifted(x, 0); }
FX_TRIG_FN float fx_cosf(float x) { return fx_trig_shifted(x, 1); }

#undef FX_TRIG_FN

#endif /* FX_TRIG_HPP */

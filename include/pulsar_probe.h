/*
 * pulsar_probe.h — device throughput probe used by the benchmark.
 *
 * No counterpart in the reference (Pulsar.jl has no device profiling). It
 * serves the benchmark's roofline: the sustained dense MFMA rate of the
 * device, reported beside the published peaks. Library: libpulsar_probe.so.
 */
#ifndef PULSAR_PROBE_H
#define PULSAR_PROBE_H

#ifdef __cplusplus
extern "C" {
#endif

#define PR_PROBE_F64 0           /* v_mfma_f64_16x16x4f64 */
#define PR_PROBE_F32 1           /* v_mfma_f32_16x16x4f32 */
#define PR_PROBE_VALU_F64 2      /* FP64 FMAs on the vector ALU */
#define PR_PROBE_F64_WITH_VALU 3 /* kinds 0 and 2 interleaved in one loop, equal
                                    issue time each: the rate shows whether FP64
                                    MFMA and FP64 VALU work overlap */

/* Runs a dense loop of the given kind on the current device (eight
 * wavefronts per SIMD, four independent MFMA accumulation chains each, no
 * memory traffic in the loop) and writes the sustained rate in TFLOP/s (for
 * kind 3: MFMA and FMA flops together). Returns 0, or a negative code:
 * -1 bad argument, -2 no device, -3 allocation, -4 launch or timing. */
int pr_probe_mfma(int kind, double* tflops);

#ifdef __cplusplus
}
#endif

#endif /* PULSAR_PROBE_H */

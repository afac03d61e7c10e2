/*
 * gen_p2ops.c -- build step: write the P2 Fekete reference-element operators as constants.
 *
 * Reads the operator files (input_files/precomputed_matrix, the files DG_leap_Frog_OpenMP.c:52-81
 * reads) with the library's own reader, forms Dr = Vr*inverV and Ds = Vs*inverV in FP64 in the same
 * order as the solver does at dgx_create, and prints them with the face lifts S1..S3 as exact hex
 * literals.  The element kernels compiled against this table take the operators as immediates instead
 * of kernel arguments: the 126 FP64 values do not fit the scalar register file, and as arguments the
 * compiler keeps part of them in vector-register lanes.  dgx_create uses these kernels only when the
 * caller's operators equal the table bit for bit.
 *
 *   gen_p2ops <operator dir>  > dgx_p2ops.h
 */
#include <stdarg.h>
#include <stdio.h>

#include "dgx_setup.h"

/* dgx_setup.c reports errors through the library's error slot; the tool prints them */
void dgx_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s <operator dir>\n", argv[0]);
        return 2;
    }
    dgx_operators ops;
    if (dgx_read_operators(argv[1], &ops) != DGX_OK) return 1;
    double v[126];
    for (int i = 0; i < 6; i++)
        for (int n = 0; n < 6; n++) {
            double dr = 0, ds = 0;
            for (int k = 0; k < 6; k++) {
                dr += ops.Vr[i * 6 + k] * ops.inverV[k * 6 + n];
                ds += ops.Vs[i * 6 + k] * ops.inverV[k * 6 + n];
            }
            v[i * 6 + n] = dr;
            v[36 + i * 6 + n] = ds;
        }
    const double *S[3] = {ops.S1, ops.S2, ops.S3};
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 18; k++) v[72 + j * 18 + k] = S[j][k];
    printf("/* Generated by gen_p2ops from %s: Dr[36] | Ds[36] | L[54] of the P2 Fekete element. */\n", argv[1]);
    printf("#ifndef DGX_P2OPS_H\n#define DGX_P2OPS_H\n\n");
    printf("static constexpr double dgx_p2_ops[126] = {\n");
    for (int k = 0; k < 126; k++) printf("    %a,%s", v[k], (k % 3 == 2 || k == 125) ? "\n" : "");
    printf("};\n\n#endif\n");
    return 0;
}

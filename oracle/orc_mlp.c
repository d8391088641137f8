/*
 * orc_mlp.c -- the learned warm start's network, restated for checking.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * No reference counterpart (the learned coarse correction is new).  MLP
 * 24 -> 128 -> 128 -> 12 with ReLU; the matrices are used rounded to BF16
 * (round to nearest even), biases in FP32; loss = mean over samples and
 * outputs of (y - t)^2.  Forward and backward in FP64 on the rounded values.
 * Parameter layout: W1[24][128] b1[128] W2[128][128] b2[128] W3[128][12] b3[12].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

enum { NI = 24, NH = 128, NO = 12 };
enum { OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
       OB3 = OW3 + NH * NO, NP = OB3 + NO };

float orc_bf16_round(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

int orc_mlp_params(void) { return NP; }

double orc_mlp_loss_grad(const float* wf, const float* X, const float* T, int n, double* grad,
                         float* Y)
{
    double* w = (double*)malloc(sizeof(double) * NP);
    for(int p = 0; p < NP; p++)
    {
        int mat = (p < OB1) || (p >= OW2 && p < OB2) || (p >= OW3 && p < OB3);
        w[p]    = mat ? (double)orc_bf16_round(wf[p]) : (double)wf[p];
    }
    if(grad)
        memset(grad, 0, sizeof(double) * NP);
    double loss = 0.0;
    double h1[NH], h2[NH], y[NO], dy[NO], d2[NH], d1[NH];
    for(int s = 0; s < n; s++)
    {
        const float* f = X + (size_t)NI * s;
        for(int h = 0; h < NH; h++)
        {
            double a = w[OB1 + h];
            for(int k = 0; k < NI; k++)
                a += (double)f[k] * w[OW1 + k * NH + h];
            h1[h] = a > 0.0 ? a : 0.0;
        }
        for(int h = 0; h < NH; h++)
        {
            double a = w[OB2 + h];
            for(int j = 0; j < NH; j++)
                a += h1[j] * w[OW2 + j * NH + h];
            h2[h] = a > 0.0 ? a : 0.0;
        }
        for(int o = 0; o < NO; o++)
        {
            double a = w[OB3 + o];
            for(int j = 0; j < NH; j++)
                a += h2[j] * w[OW3 + j * NO + o];
            y[o] = a;
            if(Y)
                Y[(size_t)NO * s + o] = (float)a;
            double e = a - (double)T[(size_t)NO * s + o];
            loss += e * e / ((double)NO * n);
            dy[o] = 2.0 * e / ((double)NO * n);
        }
        if(!grad)
            continue;
        for(int j = 0; j < NH; j++)
        {
            double g = 0.0;
            for(int o = 0; o < NO; o++)
            {
                grad[OW3 + j * NO + o] += h2[j] * dy[o];
                g += dy[o] * w[OW3 + j * NO + o];
            }
            d2[j] = h2[j] > 0.0 ? g : 0.0;
        }
        for(int o = 0; o < NO; o++)
            grad[OB3 + o] += dy[o];
        for(int j = 0; j < NH; j++)
        {
            double g = 0.0;
            for(int h = 0; h < NH; h++)
            {
                grad[OW2 + j * NH + h] += h1[j] * d2[h];
                g += d2[h] * w[OW2 + j * NH + h];
            }
            d1[j] = h1[j] > 0.0 ? g : 0.0;
        }
        for(int h = 0; h < NH; h++)
        {
            grad[OB2 + h] += d2[h];
            grad[OB1 + h] += d1[h];
            for(int k = 0; k < NI; k++)
                grad[OW1 + k * NH + h] += (double)f[k] * d1[h];
        }
    }
    free(w);
    return loss;
}

// cv_pol.h -- Lambda polarization from the freeze-out surface (pyvisc/kernel/
// kernel_polarization.cl: polarization_on_sf :14-148 (thermal vorticity), ..._shear_on_sf
// :151-338, ..._accT_on_sf :341-495; driver pyvisc/polarization/polarization.py).
// One workgroup per momentum (m_T, Y, p_x, p_y): every thread sums the elements i = tid,
// tid + 256, ... in registers, then one wave/LDS reduction (the reference reduces a tree per
// chunk of 256 elements; the sums agree to FP32 accumulation).
#pragma once
#include <hip/hip_runtime.h>

namespace cv {

struct PolP {
    int64_t nsf;
    float tfrz, beta, mass;    // beta = 1/T_frz rounded once from double (the reference's 1.0/Tfrz)
    int kind;                  // 0 thermal, 1 shear, 2 accT
};

template <int KIND>
__global__ __launch_bounds__(256) void k_polarization(PolP P, const float *__restrict__ sf, const float *__restrict__ om1,
                                                      const float *__restrict__ om2, const float *__restrict__ mom,
                                                      float *pol_lab, float *density, float *pol_lrf) {
    const int g = blockIdx.x, tid = threadIdx.x;
    const float mt = mom[4 * g], Y = mom[4 * g + 1], px = mom[4 * g + 2], py = mom[4 * g + 3];
    const float hbarc = 0.19733f;
    float acc_d = 0.0f, a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (int64_t i = tid; i < P.nsf; i += blockDim.x) {
        const float *s = sf + 8 * i;
        const float vx = s[4], vy = s[5], vz = s[6], etas = s[7];
        const float gi = 1.0f / sqrtf(fmaxf(1.0E-10f, 1.0f - vx * vx - vy * vy - vz * vz));
        const float u0 = 1.0f * gi, u1 = vx * gi, u2 = vy * gi, u3 = vz * gi;       // u^mu
        const float ch = coshf(Y - etas), sh = sinhf(Y - etas);
        const float mtcosh = mt * ch, mtsinh = mt * sh;
        const float p0 = mtcosh, p1 = -px, p2 = -py, p3 = -mtsinh;                  // p_mu
        const float pdotu = u0 * p0 + u1 * p1 + u2 * p2 + u3 * p3;
        const float tmp = expf(-pdotu * P.beta);
        const float feq = tmp / (1.0f + tmp);
        const float volum = s[0] * p0 + s[1] * p1 + s[2] * p2 + s[3] * p3;
        float otx, oty, otz, oxy, oxz, oyz, q0, q1, q2, q3;
        if (KIND == 1) {
            // symmetric gradients 2 d_(mu u_nu) from omega_shear1 (0.5 d_mu u_nu), minus p.u DU
            const float *w = om1 + 16 * i, *w2 = om2 + 4 * i;
            float sym[4][4];
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) sym[a][b] = a == b ? w[5 * a] * 2.0f : w[4 * b + a] + w[4 * a + b];
            const float pu0 = mtcosh, pu1 = px, pu2 = py, pu3 = mtsinh;               // p^mu
            float ot[4];
            for (int a = 0; a < 4; a++)
                ot[a] = (pu0 * sym[a][0] + pu1 * sym[a][1] + pu2 * sym[a][2] + pu3 * sym[a][3]) - pdotu * w2[a];
            const float l1 = -vx * gi, l2 = -vy * gi, l3 = -vz * gi, l0 = 1.0f * gi;    // u_mu
            otx = p2 * l3 - p3 * l2;
            oty = p3 * l1 - p1 * l3;
            otz = p1 * l2 - p2 * l1;
            oxy = p0 * l3 - p3 * l0;
            oxz = p2 * l0 - p0 * l2;
            oyz = p0 * l1 - p1 * l0;
            q0 = ot[0]; q1 = ot[1]; q2 = ot[2]; q3 = ot[3];
        } else {
            const float *w = om1 + 6 * i;
            otx = w[0]; oty = w[1]; otz = w[2]; oxy = w[3]; oxz = w[4]; oyz = w[5];
            q0 = p0; q1 = p1; q2 = p2; q3 = p3;
        }
        const float pw0 = otx * q1 + oty * q2 + otz * q3;
        const float pw1 = -otx * q0 + oxy * q2 + oxz * q3;
        const float pw2 = -oty * q0 - oxy * q1 + oyz * q3;
        const float pw3 = -otz * q0 - oxz * q1 - oyz * q2;
        float f;
        if (KIND == 0) f = volum * feq * (1.0f - feq);
        else if (KIND == 1) f = -hbarc * volum * feq * (1.0f - feq) / (pdotu * P.tfrz);
        else f = -hbarc * volum * feq * (1.0f - feq) / P.tfrz;
        const float s0 = f * pw0, s3 = f * pw3;
        const float ce = coshf(etas), se = sinhf(etas);
        acc_d += volum * feq;
        a0 += ce * s0 + se * s3;
        a1 += f * pw1;
        a2 += f * pw2;
        a3 += se * s0 + ce * s3;
    }
    // block reduction: wave shuffles, then the four waves through LDS
    for (int off = 32; off > 0; off >>= 1) {
        acc_d += __shfl_xor(acc_d, off, 64);
        a0 += __shfl_xor(a0, off, 64);
        a1 += __shfl_xor(a1, off, 64);
        a2 += __shfl_xor(a2, off, 64);
        a3 += __shfl_xor(a3, off, 64);
    }
    __shared__ float red[4][5];
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) { red[wv][0] = acc_d; red[wv][1] = a0; red[wv][2] = a1; red[wv][3] = a2; red[wv][4] = a3; }
    __syncthreads();
    if (tid == 0) {
        float r[5];
        for (int c = 0; c < 5; c++) r[c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
        const float dens = r[0];
        const float m4 = 4.0f * P.mass;
        density[g] = dens;
        pol_lab[4 * g + 0] = r[1] / m4; pol_lab[4 * g + 1] = r[2] / m4;
        pol_lab[4 * g + 2] = r[3] / m4; pol_lab[4 * g + 3] = r[4] / m4;
        // the local rest frame of the Lambda
        const float md = m4 * dens;
        const float L0 = r[1] / md, L1 = r[2] / md, L2 = r[3] / md, L3 = r[4] / md;
        const float e0 = mt * coshf(Y), pz = mt * sinhf(Y);
        const float q0_ = e0, q1_ = -px, q2_ = -py, q3_ = -pz;
        const float Pi0 = (q0_ * L0 + q1_ * L1 + q2_ * L2 + q3_ * L3) / P.mass;
        const float d3 = q1_ * L1 + q2_ * L2 + q3_ * L3;
        const float den = e0 * (e0 + P.mass);
        pol_lrf[4 * g + 0] = Pi0;
        pol_lrf[4 * g + 1] = L1 - d3 * q1_ / den;
        pol_lrf[4 * g + 2] = L2 - d3 * q2_ / den;
        pol_lrf[4 * g + 3] = L3 - d3 * q3_ / den;
    }
}

}  // namespace cv

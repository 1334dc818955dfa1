// lane layout probe of v_mfma_f32_16x16x16_bf16 (one wave, in-bounds)
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));
static short bf(float x) { unsigned u; memcpy(&u, &x, 4); return (short)(u >> 16); }
__global__ void k(const v4s* a, const v4s* b, v4f* d) {
    int l = threadIdx.x;
    v4f c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[l], b[l], c, 0, 0, 0);
    d[l] = c;
}
int main() {
    float A[16][16], B[16][16];
    for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) { A[i][j] = (float)((i * 16 + j) % 7 + 1); B[i][j] = (float)((i * 3 + j) % 5 + 1); }
    // assumed layout: lane (q, c) = (l / 16, l % 16): A[c][4q + k], B[4q + k][c]
    v4s ha[64], hb[64];
    for (int l = 0; l < 64; l++) {
        int q = l / 16, c = l % 16;
        for (int kk = 0; kk < 4; kk++) { ha[l][kk] = bf(A[c][4 * q + kk]); hb[l][kk] = bf(B[4 * q + kk][c]); }
    }
    v4s *da, *db; v4f* dd;
    hipMalloc(&da, sizeof(ha)); hipMalloc(&db, sizeof(hb)); hipMalloc(&dd, 64 * sizeof(v4f));
    hipMemcpy(da, ha, sizeof(ha), hipMemcpyHostToDevice);
    hipMemcpy(db, hb, sizeof(hb), hipMemcpyHostToDevice);
    k<<<1, 64>>>(da, db, dd);
    v4f hd[64];
    hipMemcpy(hd, dd, sizeof(hd), hipMemcpyDeviceToHost);
    // candidate: element r of lane (q, c) = D[4q + r][c]
    int bad1 = 0, bad2 = 0;
    for (int l = 0; l < 64; l++) {
        int q = l / 16, c = l % 16;
        for (int r = 0; r < 4; r++) {
            float ref1 = 0, ref2 = 0;
            for (int kk = 0; kk < 16; kk++) { ref1 += A[4 * q + r][kk] * B[kk][c]; ref2 += A[4 * r + q][kk] * B[kk][c]; }
            if (hd[l][r] != ref1) bad1++;
            if (hd[l][r] != ref2) bad2++;
        }
    }
    printf("layout A[c][4q+k] B[4q+k][c]: D[4q+r][c] mismatches %d, D[4r+q][c] mismatches %d\n", bad1, bad2);
    printf("lane0 %g %g %g %g lane17 %g %g %g %g\n", hd[0][0], hd[0][1], hd[0][2], hd[0][3], hd[17][0], hd[17][1], hd[17][2], hd[17][3]);
    return (bad1 == 0 || bad2 == 0) ? 0 : 1;
}

// Lane layout of gfx950's ds_read_b64_tr_b16, as the device computes it: LDS holds s[i] = i
// (1024 16-bit values); each of the 64 lanes of one wave passes an 8-byte-aligned address by one
// of three patterns, all inside the 2 KB array, and prints the four values it receives. Its
// output on an MI355X is ds_read_tr_probe.txt beside this file (tests/test_lds_transpose_model.py
// checks its model of the instruction against it). Build and run on the GPU:
//   hipcc --offload-arch=gfx950 -O2 -o probe ds_read_tr_probe.hip && ./probe > ds_read_tr_probe.txt
#include <hip/hip_runtime.h>

#include <cstdio>

typedef short shortx4 __attribute__((ext_vector_type(4)));

__global__ void probe(short *o) {
    __shared__ short s[1024];
    for (int i = threadIdx.x; i < 1024; i += 64) s[i] = static_cast<short>(i);
    __syncthreads();
    const int l = threadIdx.x;
    const int addr[3] = {4 * l, 64 * (l % 16) + 4 * (l / 16), 16 * l};
    for (int p = 0; p < 3; ++p) {
        const shortx4 v = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
            (__attribute__((address_space(3))) shortx4 *)(s + addr[p]));
        for (int j = 0; j < 4; ++j) o[(p * 64 + l) * 4 + j] = v[j];
    }
}

int main() {
    short *d = nullptr, h[3 * 64 * 4];
    if (hipMalloc(&d, sizeof h) != hipSuccess) return 1;
    hipLaunchKernelGGL(probe, dim3(1), dim3(64), 0, 0, d);
    if (hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    for (int p = 0; p < 3; ++p) {
        std::printf("pattern %d\n", p);
        for (int l = 0; l < 64; ++l)
            std::printf("%2d: %4d %4d %4d %4d\n", l, h[(p * 64 + l) * 4], h[(p * 64 + l) * 4 + 1],
                        h[(p * 64 + l) * 4 + 2], h[(p * 64 + l) * 4 + 3]);
    }
    return hipFree(d) == hipSuccess ? 0 : 3;
}

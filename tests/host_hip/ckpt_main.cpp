// Writes fields of given sizes (filled with a fixed byte pattern) to a
// checkpoint file and, with "read", reads them back and compares.  A read
// starts from buffers filled with 0x5a; when it fails, every byte must still
// be 0x5a (a refused file leaves the destination untouched), else exit 8.
#include "hip/hip_runtime.h"
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "ctsm_gpu.h"
namespace ctsm {
int ckpt_write(hipStream_t, const char *, const ctsm_ckpt_field *, int, hipError_t *);
int ckpt_read(hipStream_t, const char *, const ctsm_ckpt_field *, int, hipError_t *);
}
int main(int argc, char **argv)
{
    // argv: mode path size...
    const bool rd = argv[1][0] == 'r';
    const int n = argc - 3;
    std::vector<std::vector<unsigned char>> a(n);
    std::vector<ctsm_ckpt_field> f(n);
    std::vector<std::string> names(n);
    for (int i = 0; i < n; ++i) {
        const size_t sz = strtoull(argv[3 + i], nullptr, 10);
        a[i].resize(sz);
        for (size_t k = 0; k < sz; ++k) a[i][k] = rd ? 0x5a : (unsigned char)((k * 131 + i * 7) & 0xff);
        names[i] = "field" + std::to_string(i);
        f[i] = {names[i].c_str(), sz ? a[i].data() : nullptr, sz};
    }
    hipError_t he;
    const int rc = rd ? ctsm::ckpt_read(nullptr, argv[2], f.data(), n, &he)
                      : ctsm::ckpt_write(nullptr, argv[2], f.data(), n, &he);
    if (rd && rc == 0)
        for (int i = 0; i < n; ++i)
            for (size_t k = 0; k < a[i].size(); ++k)
                if (a[i][k] != (unsigned char)((k * 131 + i * 7) & 0xff)) { printf("mismatch %d %zu\n", i, k); return 9; }
    if (rd && rc != 0)
        for (int i = 0; i < n; ++i)
            for (size_t k = 0; k < a[i].size(); ++k)
                if (a[i][k] != 0x5a) { printf("dirty %d %zu\n", i, k); return 8; }
    printf("%d\n", rc);
    return 0;
}

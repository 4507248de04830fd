// A reference-style driver (the shape of samples/memcopy/fs.cu) built against
// gpufs_compat.h: every workgroup gopens the input file, greads one range of
// it into a device buffer and gcloses it; the host checks the bytes.
//   compat_copy <file> [workgroups]
// Exit status 0 when every byte matches the file.
#include <cstring>
#include <vector>
#include <fcntl.h>
#include <unistd.h>
#include "gpufs_compat.h"

__global__ void copy_ranges(const char* name, uint64_t range, uint8_t* dst, int64_t* res) {
    int fd = gopen(name, O_GRDONLY);
    if (fd < 0) {
        if (threadIdx.x == 0) res[blockIdx.x] = fd;
        return;
    }
    size_t got = gread(fd, (size_t)blockIdx.x * range, range, dst + (size_t)blockIdx.x * range);
    int c = gclose(fd);
    if (threadIdx.x == 0) res[blockIdx.x] = (int64_t)got < 0 ? (int64_t)got : (c ? c : (int64_t)got);
}

#define CHECK_HIP(x) do { hipError_t e = (x); if (e != hipSuccess) { fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e)); return 2; } } while (0)

int main(int argc, char** argv) {
    if (argc < 2) { fprintf(stderr, "usage: %s <file> [workgroups]\n", argv[0]); return 2; }
    const int nwg = argc > 2 ? atoi(argv[2]) : 16;
    int fd = open(argv[1], O_RDONLY);
    if (fd < 0) { perror("open"); return 2; }
    const uint64_t size = (uint64_t)lseek(fd, 0, SEEK_END);
    std::vector<uint8_t> want(size);
    if (pread(fd, want.data(), size, 0) != (ssize_t)size) { perror("pread"); return 2; }
    close(fd);
    const uint64_t range = (size + nwg - 1) / nwg;

    volatile GPUGlobals* gpuGlobals;
    initializer(&gpuGlobals);
    if (!gpuGlobals) return 2;
    char* d_name = nullptr;
    uint8_t* d_dst = nullptr;
    int64_t* d_res = nullptr;
    CHECK_HIP(hipMalloc(&d_name, strlen(argv[1]) + 1));
    CHECK_HIP(hipMemcpy(d_name, argv[1], strlen(argv[1]) + 1, hipMemcpyHostToDevice));
    CHECK_HIP(hipMalloc(&d_dst, range * nwg));
    CHECK_HIP(hipMalloc(&d_res, sizeof(int64_t) * nwg));
    copy_ranges<<<nwg, 256, 0, gpuGlobals->streamMgr->kernelStream>>>(d_name, range, d_dst, d_res);
    CHECK_HIP(hipGetLastError());
    int rc = run_gpufs_handler(gpuGlobals, 0);
    if (rc) { fprintf(stderr, "run_gpufs_handler: %s\n", gpufs_strerror(rc)); return 1; }
    std::vector<uint8_t> got(size);
    std::vector<int64_t> res(nwg);
    CHECK_HIP(hipMemcpy(got.data(), d_dst, size, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(res.data(), d_res, sizeof(int64_t) * nwg, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int i = 0; i < nwg; ++i) {
        const uint64_t lo = (uint64_t)i * range, hi = lo + range < size ? lo + range : size;
        const int64_t exp = lo < size ? (int64_t)(hi - lo) : 0;
        if (res[i] != exp) { fprintf(stderr, "workgroup %d: %lld, expected %lld\n", i, (long long)res[i], (long long)exp); bad = 1; }
    }
    if (memcmp(got.data(), want.data(), size) != 0) { fprintf(stderr, "bytes differ\n"); bad = 1; }
    (void)hipFree(d_name); (void)hipFree(d_dst); (void)hipFree(d_res);
    delete gpuGlobals;
    printf(bad ? "compat_copy FAILED\n" : "compat_copy OK (%llu bytes, %d workgroups)\n", (unsigned long long)size, nwg);
    return bad;
}

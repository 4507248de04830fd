// CPU test of the host service's descriptor cache (host_service.h DescCache):
// a reopen of an unchanged path reuses the kept descriptor and reports the
// file's current size; a path renamed over or removed takes a real open;
// flags beyond the access mode are never cached; the cap evicts the least
// recently closed descriptors; threads opening and closing the same paths
// concurrently (under ThreadSanitizer when available) never share a
// descriptor.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>
#include <dirent.h>
#include "../../gpufs_amd/csrc/host_service.h"

static int failures = 0;
#define CHECK(c, ...)                                                              \
    do {                                                                           \
        if (!(c)) {                                                                \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                   \
            fprintf(stderr, __VA_ARGS__);                                          \
            fprintf(stderr, "\n");                                                 \
            ++failures;                                                            \
        }                                                                          \
    } while (0)

static int open_fds() {
    int n = 0;
    DIR* d = opendir("/proc/self/fd");
    if (!d) return -1;
    while (readdir(d)) ++n;
    closedir(d);
    return n;
}
static void put(const std::string& p, size_t n, char c) {
    FILE* f = fopen(p.c_str(), "wb");
    std::string s(n, c);
    fwrite(s.data(), 1, n, f);
    fclose(f);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    const std::string a = dir + "/dc_a.bin", b = dir + "/dc_b.bin", tmp = dir + "/dc_tmp.bin";
    put(a, 1000, 'a');
    put(b, 2000, 'b');
    const int base = open_fds();
    {
        gfs_host::DescCache dc;
        struct stat st;
        // first open: a real one, kept after the close
        int64_t f1 = dc.open(a.c_str(), O_RDONLY, &st);
        CHECK(f1 >= 0 && st.st_size == 1000, "open a: %lld size %lld", (long long)f1, (long long)st.st_size);
        CHECK(dc.close((int)f1) == 0, "close a");
        CHECK(dc.size() == 1, "one kept");
        // reopen: the same descriptor, and the size of the file now
        put(a, 1500, 'A');   // same inode, new size and times
        int64_t f2 = dc.open(a.c_str(), O_RDONLY, &st);
        CHECK(f2 == f1, "reopen reuses %lld vs %lld", (long long)f2, (long long)f1);
        CHECK(st.st_size == 1500, "reopen sees size %lld", (long long)st.st_size);
        // a second open while the kept one is in use: a real open
        struct stat st2;
        int64_t f3 = dc.open(a.c_str(), O_RDONLY, &st2);
        CHECK(f3 >= 0 && f3 != f2, "busy: another descriptor (%lld)", (long long)f3);
        CHECK(dc.close((int)f3) == 0 && dc.close((int)f2) == 0, "closes");
        // another access mode is another entry
        int64_t w = dc.open(a.c_str(), O_RDWR, &st);
        CHECK(w >= 0 && w != f1, "rdwr is separate");
        CHECK(dc.close((int)w) == 0, "close rdwr");
        // renamed over: the reopen finds another inode and opens for real
        put(tmp, 300, 't');
        CHECK(rename(tmp.c_str(), a.c_str()) == 0, "rename");
        int64_t f4 = dc.open(a.c_str(), O_RDONLY, &st);
        CHECK(f4 >= 0 && st.st_size == 300, "after rename: size %lld", (long long)st.st_size);
        char c = 0;
        CHECK(pread((int)f4, &c, 1, 0) == 1 && c == 't', "after rename reads the new file");
        CHECK(dc.close((int)f4) == 0, "close");
        // removed: the kept descriptor goes, the open fails as a real one would
        CHECK(unlink(a.c_str()) == 0, "unlink");
        int64_t f5 = dc.open(a.c_str(), O_RDONLY, &st);
        CHECK(f5 == -ENOENT, "removed file: %lld", (long long)f5);
        // O_CREAT / O_TRUNC are never kept
        int64_t f6 = dc.open(a.c_str(), O_RDWR | O_CREAT, &st);
        CHECK(f6 >= 0, "create");
        const size_t kept = dc.size();
        CHECK(dc.close((int)f6) == 0, "close created");
        CHECK(dc.size() == kept, "created not kept");
        // threads on the same few paths: a descriptor is never handed out twice
        std::vector<std::thread> th;
        std::atomic<int> bad{0};
        std::vector<std::atomic<int>> owner(4096);
        for (auto& o : owner) o = 0;
        for (int t = 0; t < 8; ++t)
            th.emplace_back([&, t] {
                struct stat s;
                for (int k = 0; k < 2000; ++k) {
                    const std::string& p = (k + t) % 2 ? b : a;
                    int64_t fd = dc.open(p.c_str(), O_RDONLY, &s);
                    if (fd < 0 || fd >= 4096) { bad++; continue; }
                    if (owner[fd].exchange(1) != 0) bad++;
                    owner[fd] = 0;
                    if (dc.close((int)fd) != 0) bad++;
                }
            });
        for (auto& x : th) x.join();
        CHECK(bad == 0, "%d bad opens or shared descriptors", bad.load());
    }
    CHECK(open_fds() == base, "descriptors left open: %d vs %d", open_fds(), base);
    if (failures) {
        printf("desc cache model FAILED (%d)\n", failures);
        return 1;
    }
    printf("desc cache model OK\n");
    return 0;
}

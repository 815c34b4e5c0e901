// Sanitizer driver of the CPU oracle. TEST INFRASTRUCTURE ONLY.
//
// Built with AddressSanitizer and UndefinedBehaviorSanitizer as one
// executable together with the oracle's source (the oracle is included as a
// translation unit, so the sanitizers see every access it makes), it runs the
// whole detector in both modes on binary PGM (P5) images and prints, per
// image, mode and decimation factor, the quad count, the stage counts and an
// FNV-1a hash of the quads (corners, reversed_border and blob_index, no
// padding) so that a test can compare them with the in-process oracle.
//
//   oracle_san IMAGE.pgm [IMAGE.pgm ...]
#include "atq_oracle.cpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

bool read_pgm(const char *path, std::vector<uint8_t> &px, int64_t &w, int64_t &h) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char magic[3] = {0, 0, 0};
    int maxv = 0;
    long lw = 0, lh = 0;
    bool ok = std::fscanf(f, "%2s %ld %ld %d", magic, &lw, &lh, &maxv) == 4 && magic[0] == 'P' &&
              magic[1] == '5' && maxv == 255 && lw > 0 && lh > 0 && std::fgetc(f) != EOF;
    if (ok) {
        w = lw;
        h = lh;
        px.resize((size_t)(w * h));
        ok = std::fread(px.data(), 1, px.size(), f) == px.size();
    }
    std::fclose(f);
    return ok;
}

uint64_t fnv1a(uint64_t hsh, const void *p, size_t n) {
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; i++) hsh = (hsh ^ b[i]) * 0x100000001b3ull;
    return hsh;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s IMAGE.pgm [IMAGE.pgm ...]\n", argv[0]);
        return 2;
    }
    for (int a = 1; a < argc; a++) {
        std::vector<uint8_t> img;
        int64_t w = 0, h = 0;
        if (!read_pgm(argv[a], img, w, h)) {
            std::fprintf(stderr, "cannot read %s\n", argv[a]);
            return 2;
        }
        for (int fixed = 0; fixed < 2; fixed++) {
            for (int f = 1; f <= 2; f++) {
                std::vector<QuadCorners> q(1 << 16);
                orc_stats st{};
                int64_t n = (fixed ? orc_fx_detect : orc_detect)(img.data(), w, h, f, 1, q.data(),
                                                                 (int64_t)q.size(), &st);
                uint64_t hsh = 0xcbf29ce484222325ull;
                for (int64_t i = 0; i < std::min<int64_t>(n, (int64_t)q.size()); i++) {
                    hsh = fnv1a(hsh, q[i].corners, sizeof(q[i].corners));
                    hsh = fnv1a(hsh, &q[i].reversed_border, 1);
                    hsh = fnv1a(hsh, &q[i].blob_index, 4);
                }
                std::printf("%s %s f%d quads %lld clusters %lld points %lld hash %016llx\n",
                            argv[a], fixed ? "fixed" : "compat", f, (long long)n,
                            (long long)st.n_clusters, (long long)st.n_cluster_points,
                            (unsigned long long)hsh);
            }
        }
    }
    return 0;
}

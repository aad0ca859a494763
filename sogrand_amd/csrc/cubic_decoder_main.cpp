// Drop-in replacement for the reference decoder executable:
//
//   cubic_decoder <input_llr_file> <output_file> [options]
//
// Same positional arguments, file formats and summary lines as
// SOGRAND_C/cubic_decoder_final.c:62-314 and SOGRAND_CUDA/cubic_decoder1.cu:927-1183,
// so it slots into cubic_flow.sh unchanged. Options (after the two paths):
//   --tmax N            guess budget per component decode (default: unlimited)
//   --precision P       fp64 (default, bit-exact with SOGRAND_C) | fp32
//   --backend B         auto (default: plain, or persistent by batch size) | persistent |
//                       plain | lanegroup (fp64, final variant: the shortest time for
//                       files of up to one codeword per CU)
//   --epsilon E         EPSILON of the soft output, e.g. 2.220446049250313e-16 for the
//                       reference's cubic_decoder_updated.c (default 1e-40)
//   --variant V         final (default: cubic_decoder_final.c) | legacy (the earlier
//                       cubic_decoder.c, whose summary starts with one more line,
//                       cubic_decoder.c:104)
//   --stats FILE        per-codeword statistics, 24-byte records
//                       {u32 half_iters, u32 converged, u64 NG, u64 NG_p}
//   --soft-out FILE     final L_APP, 4096 float64 per codeword
//   --device N          HIP device ordinal (default 0)
//   --dim D             3 (default: the cubic code) | 2 (the (16,8)^2 square code:
//                       256 float64 LLRs in and 8 info bytes out per codeword, the
//                       dimension SOGRAND_C/channel_sim.c also accepts)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sogrand.h"

static int usage(const char *argv0)
{
    std::fprintf(stderr, "Usage: %s <input_llr_file> <output_file>\n", argv0);
    return 1;
}

int main(int argc, char *argv[])
{
    if (argc < 3) return usage(argv[0]);
    const char *in_path = argv[1];
    const char *out_path = argv[2];
    sogrand_params p;
    sogrand_params_default(&p);
    p.backend = SOGRAND_BACKEND_AUTO; // files of any size: the kernel that suits it
    const char *stats_path = nullptr, *soft_path = nullptr;
    int device = 0, dim = 3;
    for (int i = 3; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](void) -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--tmax") {
            p.Tmax = std::strtoull(need(), nullptr, 10);
        } else if (a == "--precision") {
            std::string v = need();
            if (v == "fp64") p.precision = SOGRAND_FP64;
            else if (v == "fp32") p.precision = SOGRAND_FP32_MIXED;
            else return usage(argv[0]);
        } else if (a == "--backend") {
            std::string v = need();
            if (v == "plain") p.backend = SOGRAND_BACKEND_PLAIN;
            else if (v == "persistent") p.backend = SOGRAND_BACKEND_PERSISTENT;
            else if (v == "lanegroup") p.backend = SOGRAND_BACKEND_LANEGROUP;
            else if (v == "auto") p.backend = SOGRAND_BACKEND_AUTO;
            else return usage(argv[0]);
        } else if (a == "--epsilon") {
            p.epsilon = std::strtod(need(), nullptr);
        } else if (a == "--variant") {
            std::string v = need();
            if (v == "final") p.variant = SOGRAND_VARIANT_FINAL;
            else if (v == "legacy") p.variant = SOGRAND_VARIANT_LEGACY;
            else return usage(argv[0]);
        } else if (a == "--stats") {
            stats_path = need();
        } else if (a == "--soft-out") {
            soft_path = need();
        } else if (a == "--device") {
            device = std::atoi(need());
        } else if (a == "--dim") {
            dim = std::atoi(need());
            if (dim != 2 && dim != 3) return usage(argv[0]);
        } else {
            return usage(argv[0]);
        }
    }

    FILE *fin = std::fopen(in_path, "rb");
    if (!fin) {
        std::perror("Error opening input file");
        return 1;
    }
    std::fseek(fin, 0, SEEK_END);
    const long long file_size = std::ftell(fin);
    std::rewind(fin);
    // positions and info bytes of one codeword of the code
    const size_t npos = dim == 3 ? SOGRAND_CUBE : 256, ninfo = dim == 3 ? SOGRAND_INFO_BYTES : 8;
    // a trailing partial codeword is ignored (cubic_decoder1.cu:997, final.c:119)
    const size_t total = (size_t)(file_size / (long long)(npos * sizeof(double)));
    if (total == 0) {
        std::fprintf(stderr, "Error: No valid blocks found in input file\n");
        std::fclose(fin);
        return 1;
    }
    FILE *fout = std::fopen(out_path, "wb");
    if (!fout) {
        std::perror("Error opening output file");
        std::fclose(fin);
        return 1;
    }
    FILE *fstats = stats_path ? std::fopen(stats_path, "wb") : nullptr;
    FILE *fsoft = soft_path ? std::fopen(soft_path, "wb") : nullptr;
    if ((stats_path && !fstats) || (soft_path && !fsoft)) {
        std::perror("Error opening statistics/soft output file");
        return 1;
    }

    sogrand_ctx *ctx = nullptr;
    int rc = sogrand_ctx_create(&ctx, device, &p);
    if (rc) {
        std::fprintf(stderr, "Error: %s\n", sogrand_strerror(rc));
        return 1;
    }
    if (p.variant == SOGRAND_VARIANT_LEGACY)
        std::printf("Decoding %s to %s using %s code (n=%d, k=%d)...\n", in_path, out_path,
                    dim == 3 ? "cubic" : "square", SOGRAND_N, SOGRAND_K);
    // chunks of the file alternate between two buffers: the next one is read by
    // a second thread while the current one decodes
    const size_t chunk = 65536;
    std::vector<double> llr[2];
    bool read_ok[2] = {false, false};
    auto read_chunk = [&](int b, size_t n) {
        llr[b].resize(n * npos);
        read_ok[b] = std::fread(llr[b].data(), sizeof(double), n * npos, fin) == n * npos;
    };
    std::vector<uint8_t> bits;
    std::vector<sogrand_cw_stats> st;
    std::vector<double> soft;
    uint64_t total_ng = 0, total_ngp = 0, total_half = 0;
    double gpu_ms = 0.0;
    read_chunk(0, total < chunk ? total : chunk);
    int cur = 0;
    for (size_t off = 0; off < total; off += chunk, cur ^= 1) {
        const size_t n = (total - off) < chunk ? (total - off) : chunk;
        std::thread reader;
        if (off + chunk < total) {
            const size_t left = total - off - chunk;
            reader = std::thread(read_chunk, cur ^ 1, left < chunk ? left : chunk);
        }
        bits.resize(n * ninfo);
        st.resize(n);
        if (fsoft) soft.resize(n * npos);
        if (!read_ok[cur]) {
            if (reader.joinable()) reader.join();
            std::fprintf(stderr, "Error: Failed to read entire file\n");
            return 1;
        }
        rc = (dim == 3 ? sogrand_decode_cubic : sogrand_decode_square)(
            ctx, llr[cur].data(), n, bits.data(), fsoft ? soft.data() : nullptr, st.data(),
            nullptr);
        if (reader.joinable()) reader.join();
        if (rc) {
            std::fprintf(stderr, "Error: %s\n", sogrand_strerror(rc));
            return 1;
        }
        double ms = 0.0;
        sogrand_last_kernel_ms(ctx, &ms, nullptr);
        gpu_ms += ms;
        for (size_t i = 0; i < n; ++i) {
            total_ng += st[i].NG;
            total_ngp += st[i].NG_p;
            total_half += st[i].half_iters;
        }
        std::fwrite(bits.data(), 1, bits.size(), fout);
        if (fstats) std::fwrite(st.data(), sizeof(sogrand_cw_stats), n, fstats);
        if (fsoft) std::fwrite(soft.data(), sizeof(double), soft.size(), fsoft);
    }
    std::fclose(fin);
    std::fclose(fout);
    if (fstats) std::fclose(fstats);
    if (fsoft) std::fclose(fsoft);
    sogrand_ctx_destroy(ctx);

    const double blocks = (double)total;
    const double gpu_s = gpu_ms / 1000.0;
    std::printf("Decoding complete. %zu block(s) decoded.\n", total);
    std::printf("Average iterations per block: %.2f\n", (double)total_half / 2.0 / blocks);
    std::printf("Average NG per block: %.2f\n", (double)total_ng / blocks);
    std::printf("Average NG per info bit: %.2f\n",
                (double)total_ng / (blocks * (dim == 3 ? 512.0 : 64.0)));
    std::printf("Average NG_p per block: %.2f\n", (double)total_ngp / blocks);
    std::printf("Total GPU time: %.2f seconds\n", gpu_s);
    std::printf("Throughput: %.2f blocks/sec\n", gpu_s > 0 ? blocks / gpu_s : 0.0);
    return 0;
}

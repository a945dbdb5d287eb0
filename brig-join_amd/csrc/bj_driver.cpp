// bj_driver.cpp -- command line front end; keeps the reference driver's surface.
//
// Same positional arguments and --flags as src/driver.cpp:43-198:
//   bj_driver rPath sPath outPath R S sanityCheck [chunk_size] [use_bf] [use_nvlink]
//             [use_warp_compact] [true_false_ratio] [s_chunks] [kernel_fusion]
//             [redist_skew_threshold]
//   --num_bits N --num_hashes K --partition_rs_split 0|1 --partition_size N
//   --second_partition_size N --fused_probe_breakdown 0|1 --read_val 0|1
//   --r_shuffle_enable 0|1 --r_replicate_ratio X --answer_scale X
// plus --radix_bits N --bf_bits_per_key X --val_mode 0|1 --device N.
// So the README command line of the reference parses unchanged.  Unlike the
// reference, outPath is written (output.bin, sorted rows; see bj_api.h).
#include "bj_api.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    bj_main_args a;
    bj_main_args_default(&a);
    std::vector<char*> pos;
    pos.push_back(argv[0]);
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        if (s.rfind("--", 0) == 0) {
            std::string key = s.substr(2), val;
            const size_t eq = key.find('=');
            if (eq != std::string::npos) {
                val = key.substr(eq + 1);
                key = key.substr(0, eq);
            } else if (i + 1 < argc) {
                val = argv[++i];
            }
            const char* v = val.c_str();
            if (key == "num_bits") a.override_num_bits = std::atoll(v);
            else if (key == "num_hashes") a.override_num_hashes = std::atoi(v);
            else if (key == "partition_rs_split") a.override_partition_rs_split = std::atoi(v);
            else if (key == "partition_size") a.override_partition_size = std::atoi(v);
            else if (key == "second_partition_size") a.override_second_partition_size = std::atoi(v);
            else if (key == "fused_probe_breakdown") a.fused_probe_breakdown = val.empty() ? 1 : std::atoi(v);
            else if (key == "read_val") a.read_val = val.empty() ? 1 : std::atoi(v);
            else if (key == "r_shuffle_enable") a.r_shuffle_enable = val.empty() ? 1 : std::atoi(v);
            else if (key == "r_replicate_ratio") a.r_replicate_ratio = std::atof(v);
            else if (key == "answer_scale") a.answer_scale = std::atof(v);
            else if (key == "radix_bits") a.radix_bits = std::atoi(v);
            else if (key == "bf_bits_per_key") a.bf_bits_per_key = std::atof(v);
            else if (key == "val_mode") a.val_mode = std::atoi(v);
            else if (key == "device") a.device = std::atoi(v);
            else std::fprintf(stderr, "Warning: unknown option --%s (ignored)\n", key.c_str());
        } else {
            pos.push_back(argv[i]);
        }
    }
    const int n = (int)pos.size();
    if (n < 7) {
        std::printf("./bj_driver rPath sPath outPath R S sanityCheck [chunk_size] [use_bf] [use_nvlink] "
                    "[use_warp_compact] [true_false_ratio] [s_chunks] [kernel_fusion] "
                    "[redist_skew_threshold]\n"
                    "  options: [--num_bits N] [--num_hashes K] [--partition_size N] "
                    "[--second_partition_size N] [--radix_bits N] [--bf_bits_per_key X] [--read_val 0|1] "
                    "[--answer_scale X] [--val_mode 0|1] [--device N]\n");
        return 2;
    }
    a.r_path = pos[1];
    a.s_path = pos[2];
    a.out_path = pos[3];
    a.r_size = std::atoll(pos[4]);
    a.s_size = std::atoll(pos[5]);
    a.sanity_check = std::atoi(pos[6]);
    if (n >= 8) a.chunk_size = std::atoi(pos[7]);
    if (n >= 9) a.use_bf = std::atoi(pos[8]);
    if (n >= 10) a.use_nvlink = std::atoi(pos[9]);
    if (n >= 11) a.use_warp_compact = std::atoi(pos[10]);
    if (n >= 12) a.true_false_ratio = std::atof(pos[11]);
    if (n >= 13) {
        a.s_chunks = std::atoi(pos[12]);
        if (a.s_chunks <= 0) {
            std::fprintf(stderr, "Error: s_chunks must be >= 1 (got %d)\n", a.s_chunks);
            return 2;
        }
    }
    if (n >= 14) a.kernel_fusion = std::atoi(pos[13]);
    if (n >= 15) a.redist_skew_threshold = std::atof(pos[14]);

    std::printf("=====================================\n");
    std::printf("             Join Tables             \n");
    std::printf("=====================================\n");
    std::printf("The size of tableA: %lld\nThe size of tableB: %lld\n", (long long)a.r_size,
                (long long)a.s_size);
    std::printf("Sanity Check(0: false, 1: true): %d\n", a.sanity_check);
    std::printf("BF filtering is %s\n", a.use_bf ? "enabled" : "disabled");
    std::printf("=====================================\n\n");

    bj_result res;
    const int rc = bj_gpu_main(&a, &res);
    if (rc != BJ_OK) {
        std::fprintf(stderr, "bj_gpu_main failed: %s (%s)\n", bj_strerror(rc), bj_last_error());
        return 1;
    }
    return 0;
}

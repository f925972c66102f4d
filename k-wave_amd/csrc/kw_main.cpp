// kspaceFirstOrder-HIP: drop-in for the kspaceFirstOrder-OMP / -CUDA
// executables that kspaceFirstOrder2DC/3DC/2DG/3DG run through system()
// (kspaceFirstOrder3DC.m:544).  Accepts the command line those wrappers build
// (kspaceFirstOrder3DC.m:308-524), reads the HDF5 input file, runs the time
// loop on the GPU and writes the HDF5 output file the wrapper reads back.
// Exit status 0 on success; non-zero with a message on stderr otherwise.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "kwave_io.h"

static void usage(FILE* out) {
  fprintf(out,
          "usage: kspaceFirstOrder-HIP -i <input.h5> -o <output.h5> [options]\n"
          "  -t <n>              host threads (accepted; the GPU does the work)\n"
          "  -g <device>         HIP device ordinal\n"
          "  -s <index>          first time step to record (1-based)\n"
          "  -c <level>          deflate level of the output datasets\n"
          "  --verbose <0..2>\n"
          "  --p_raw --p_max --p_min --p_rms --p_max_all --p_min_all --p_final\n"
          "  --u_raw --u_max --u_min --u_rms --u_max_all --u_min_all --u_final --u_non_staggered_raw\n"
          "  (no output flag: --p_raw)\n"
          "  --checkpoint_file <f> [--checkpoint_interval <s>] [--checkpoint_timesteps <n>]\n"
          "                      stop and save after <s> seconds / <n> steps; run again to continue\n");
}

int main(int argc, char** argv) {
  std::string in, out, ckpt_file;
  kw_run_options o;
  memset(&o, 0, sizeof o);
  o.record_start_index = 1;
  struct Flag {
    const char* name;
    uint64_t bit;
  } flags[] = {{"--p_raw", KW_REC_P},         {"--p_max", KW_REC_P_MAX},         {"--p_min", KW_REC_P_MIN},
               {"--p_rms", KW_REC_P_RMS},     {"--p_max_all", KW_REC_P_MAX_ALL}, {"--p_min_all", KW_REC_P_MIN_ALL},
               {"--p_final", KW_REC_P_FINAL}, {"--u_raw", KW_REC_U},             {"--u_max", KW_REC_U_MAX},
               {"--u_min", KW_REC_U_MIN},     {"--u_rms", KW_REC_U_RMS},         {"--u_max_all", KW_REC_U_MAX_ALL},
               {"--u_min_all", KW_REC_U_MIN_ALL}, {"--u_final", KW_REC_U_FINAL},
               {"--u_non_staggered_raw", KW_REC_U_NON_STAGGERED}};
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&](const char* what) -> const char* {
      if (i + 1 >= argc) {
        fprintf(stderr, "kspaceFirstOrder-HIP: %s needs a value\n", what);
        exit(2);
      }
      return argv[++i];
    };
    bool matched = false;
    for (const Flag& f : flags)
      if (a == f.name) {
        o.record_flags |= f.bit;
        matched = true;
      }
    if (matched) continue;
    if (a == "-i") {
      in = val("-i");
    } else if (a == "-o") {
      out = val("-o");
    } else if (a == "-t") {
      (void)val("-t");
    } else if (a == "-g") {
      o.device = atoi(val("-g"));
    } else if (a == "-s") {
      o.record_start_index = strtoull(val("-s"), nullptr, 10);
    } else if (a == "-c") {
      o.compression_level = atoi(val("-c"));
    } else if (a == "--verbose") {
      o.verbose = atoi(val("--verbose"));
    } else if (a == "-h" || a == "--help") {
      usage(stdout);
      return 0;
    } else if (a == "--version") {
      printf("kspaceFirstOrder-HIP (gfx950), HDF5 file format 1.2\n");
      return 0;
    } else if (a == "--checkpoint_file") {
      ckpt_file = val("--checkpoint_file");
    } else if (a == "--checkpoint_interval") {
      o.checkpoint_interval = atof(val("--checkpoint_interval"));
    } else if (a == "--checkpoint_timesteps") {
      o.checkpoint_timesteps = strtoull(val("--checkpoint_timesteps"), nullptr, 10);
    } else if (a == "--copy_sensor_mask" || a == "--benchmark") {
      fprintf(stderr, "kspaceFirstOrder-HIP: option %s is not supported by this build\n", a.c_str());
      return 3;
    } else {
      fprintf(stderr, "kspaceFirstOrder-HIP: unknown option %s\n", a.c_str());
      usage(stderr);
      return 2;
    }
  }
  if (in.empty() || out.empty()) {
    usage(stderr);
    return 2;
  }
  if (o.record_start_index == 0) o.record_start_index = 1;
  printf("┌───────────────────────────────────────────────────────────────┐\n");
  printf("│              kspaceFirstOrder-HIP  (MI355X, gfx950)           │\n");
  printf("└───────────────────────────────────────────────────────────────┘\n");
  printf("  input  : %s\n  output : %s\n", in.c_str(), out.c_str());
  fflush(stdout);
  if (!ckpt_file.empty()) o.checkpoint_file = ckpt_file.c_str();
  kw_status st = kw_run_file(in.c_str(), out.c_str(), &o);
  if (st != KW_OK) {
    fprintf(stderr, "kspaceFirstOrder-HIP: error %d: %s\n", (int)st, kwio_last_error());
    return 1;
  }
  if (o.checkpointed) printf("  checkpoint saved to %s; run the same command again to continue\n", ckpt_file.c_str());
  printf("  data loading    : %8.3f s\n", o.t_load);
  printf("  pre-processing  : %8.3f s\n", o.t_precompute);
  printf("  time loop       : %8.3f s\n", o.t_loop);
  printf("  output saving   : %8.3f s\n", o.t_save);
  printf("  total           : %8.3f s\n", o.t_load + o.t_precompute + o.t_loop + o.t_save);
  return 0;
}

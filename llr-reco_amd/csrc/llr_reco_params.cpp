// llr-reco-params: writes the default LlrReconParams TOML file
// (bin/llr-reco-params.rs:7-23), through the library's C ABI.
//
//   llr-reco-params <parameter_file>
#include <cstdio>
#include <string>

#include "../../include/llr_hip.h"

int main(int argc, char** argv) {
  if (argc != 2 || std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help") {
    std::fprintf(argc == 2 ? stdout : stderr, "usage: llr-reco-params <parameter_file>\n");
    return argc == 2 ? 0 : 2;
  }
  llr_toml_params p;
  llr_recon_params_default(&p);
  if (llr_recon_params_write(argv[1], &p) != LLR_OK) {
    std::fprintf(stderr, "llr-reco-params: %s\n", llr_last_error());
    return 1;
  }
  std::string f = argv[1];
  const size_t slash = f.find_last_of('/'), dot = f.find_last_of('.');
  if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) f = f.substr(0, dot);
  std::printf("wrote default parameter file to %s.toml\n", f.c_str());
  return 0;
}

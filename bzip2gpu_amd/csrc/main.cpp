// main.cpp - bzip2gpu CLI (reference: src/main.cpp:20-60).  Input is cut
// into 900000-byte chunks, each one block, at every level (the reference's
// cut); the level sets the header digit.
#include <fstream>
#include <iostream>
#include <vector>

#include "cli.h"
#include "compression.h"
#include "io.h"

static bool write_file(const std::string &name, const std::vector<uint8_t> &data) {
  std::ofstream f(name, std::ios::out | std::ios::binary);
  if (!f) {
    std::cerr << "Failed to open file for writing: " << name << std::endl;
    return false;
  }
  f.write(reinterpret_cast<const char *>(data.data()), (std::streamsize)data.size());
  return f.good();
}

int main(int argc, char **argv) {
  CliOptions options = parse_args(argc, argv);
  int rc = 0;
  if (options.input_files.empty()) {
    BZFileInputStream input(10, 900000);
    std::vector<uint8_t> output;
    bzip2_gpu_compress(input, options.block_size, output);
    std::cout.write(reinterpret_cast<const char *>(output.data()), (std::streamsize)output.size());
  } else {
    for (const auto &file : options.input_files) {
      BZFileInputStream input(10, 900000, file);
      if (input.failed) {
        std::cerr << "bzip2gpu: can't open input file " << file << std::endl;
        rc = 1;
        continue;
      }
      std::vector<uint8_t> output;
      bzip2_gpu_compress(input, options.block_size, output);
      if (options.stdout_output)
        std::cout.write(reinterpret_cast<const char *>(output.data()), (std::streamsize)output.size());
      else if (!write_file(file + ".bz2", output))
        rc = 1;
    }
  }
  return rc;
}

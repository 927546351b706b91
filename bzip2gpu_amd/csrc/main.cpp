// main.cpp - bzip2gpu CLI (reference: src/main.cpp:20-60).  Compression cuts
// the input into 900000-byte chunks, each one block, at every level (the
// reference's cut); the level sets the header digit.  Decompression (-d, or
// a program name containing "bunzip"/"bzcat") and -t run the device
// decoder; the reference has none (TODO.md:28-34).  File handling is the
// reference's: inputs are kept and outputs overwritten (bzip2's -k -f).
// Exit status as bzip2: 0 ok, 1 file problem, 2 corrupt compressed input,
// 3 internal error.
#include <stdio.h>

#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "cli.h"
#include "compression.h"
#include "io.h"

static bool read_all(std::istream &in, std::vector<uint8_t> &data) {
  data.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  return !in.bad();
}

static bool write_file(const std::string &name, const std::vector<uint8_t> &data) {
  std::ofstream f(name, std::ios::out | std::ios::binary | std::ios::trunc);
  if (!f) {
    std::cerr << "bzip2gpu: Can't create output file " << name << std::endl;
    return false;
  }
  f.write(reinterpret_cast<const char *>(data.data()), (std::streamsize)data.size());
  return f.good();
}

static void write_stdout(const std::vector<uint8_t> &data) {
  std::cout.write(reinterpret_cast<const char *>(data.data()), (std::streamsize)data.size());
  std::cout.flush();
}

static bool ends_with(const std::string &s, const std::string &x) {
  return s.size() > x.size() && s.compare(s.size() - x.size(), x.size(), x) == 0;
}

// bzip2's output names for decompression
static std::string decompressed_name(const std::string &in, bool quiet) {
  if (ends_with(in, ".bz2")) return in.substr(0, in.size() - 4);
  if (ends_with(in, ".bz")) return in.substr(0, in.size() - 3);
  if (ends_with(in, ".tbz2")) return in.substr(0, in.size() - 5) + ".tar";
  if (ends_with(in, ".tbz")) return in.substr(0, in.size() - 4) + ".tar";
  if (!quiet) std::cerr << "bzip2gpu: Can't guess original name for " << in << " -- using " << in << ".out" << std::endl;
  return in + ".out";
}

static void report_compress(const std::string &name, size_t in, size_t out) {
  if (in == 0) {
    fprintf(stderr, "  %s: no data compressed.\n", name.c_str());
    return;
  }
  const double r = (double)in / (double)out;
  fprintf(stderr, "  %s: %6.3f:1, %6.3f bits/byte, %5.2f%% saved, %zu in, %zu out.\n", name.c_str(), r,
          8.0 * (double)out / (double)in, 100.0 * (1.0 - (double)out / (double)in), in, out);
}

// Decodes (or tests) one input; returns the exit severity.
static int decompress_one(const std::string &label, const std::vector<uint8_t> &comp, const CliOptions &o,
                          const std::string *out_name) {
  std::vector<uint8_t> raw;
  std::string err;
  const int st = bzip2_gpu_decompress(comp.data(), comp.size(), raw, &err);
  if (st != 0) {
    std::cerr << "bzip2gpu: " << label << ": " << (st == 2 ? "data integrity error: " : "internal error: ") << err
              << std::endl;
    return st;
  }
  if (o.test) {
    if (o.verbose) fprintf(stderr, "  %s: ok\n", label.c_str());
    return 0;
  }
  if (o.stdout_output || !out_name) write_stdout(raw);
  else if (!write_file(*out_name, raw)) return 1;
  if (o.verbose) fprintf(stderr, "  %s: done\n", label.c_str());
  return 0;
}

int main(int argc, char **argv) {
  CliOptions options = parse_args(argc, argv);
  int rc = 0;
  auto worst = [&rc](int v) { rc = v > rc ? v : rc; };
  if (options.input_files.empty()) {
    if (options.compress) {
      BZFileInputStream input(10, 900000);
      std::vector<uint8_t> output;
      bzip2_gpu_compress(input, options.block_size, output);
      write_stdout(output);
    } else {
      std::vector<uint8_t> comp;
      if (!read_all(std::cin, comp)) {
        std::cerr << "bzip2gpu: I/O error reading standard input" << std::endl;
        return 1;
      }
      worst(decompress_one("(stdin)", comp, options, nullptr));
    }
    return rc;
  }
  for (const auto &file : options.input_files) {
    if (options.compress) {
      BZFileInputStream input(10, 900000, file);
      if (input.failed) {
        std::cerr << "bzip2gpu: Can't open input file " << file << std::endl;
        worst(1);
        continue;
      }
      std::vector<uint8_t> output;
      bzip2_gpu_compress(input, options.block_size, output);
      if (options.verbose) {
        std::ifstream f(file, std::ios::binary | std::ios::ate);
        report_compress(file, f ? (size_t)f.tellg() : 0, output.size());
      }
      if (options.stdout_output) write_stdout(output);
      else if (!write_file(file + ".bz2", output)) worst(1);
    } else {
      std::ifstream f(file, std::ios::binary);
      std::vector<uint8_t> comp;
      if (!f || !read_all(f, comp)) {
        std::cerr << "bzip2gpu: Can't open input file " << file << std::endl;
        worst(1);
        continue;
      }
      const std::string out = options.test || options.stdout_output ? std::string() : decompressed_name(file, options.quiet);
      worst(decompress_one(file, comp, options, out.empty() ? nullptr : &out));
    }
  }
  return rc;
}

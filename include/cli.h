// cli.h - drop-in for the reference's option parsing
// (SebastianBoteroLeonik/bzip2gpu include/cli.h:7-14, src/utils/argsparce.cpp).
#ifndef BZG_CLI_H
#define BZG_CLI_H

#include <string>
#include <vector>

struct CliOptions {
  bool compress = true;
  bool stdout_output = false;
  int block_size = 9;
  std::vector<std::string> input_files;
};

CliOptions parse_args(int argc, char **argv);

#endif  // BZG_CLI_H

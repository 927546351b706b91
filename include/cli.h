// cli.h - drop-in for the reference's option parsing
// (SebastianBoteroLeonik/bzip2gpu include/cli.h:7-14, src/utils/argsparce.cpp),
// with the bzip2 flags its help text lists as not yet available
// (argsparce.cpp:11-21): -d -k -f -t -q -v -L -V -s.
#ifndef BZG_CLI_H
#define BZG_CLI_H

#include <string>
#include <vector>

struct CliOptions {
  bool compress = true;        // -z (default) / -d
  bool stdout_output = false;  // -c
  int block_size = 9;          // -1 .. -9
  bool test = false;           // -t: check compressed files, write nothing
  bool keep = false;           // -k (inputs are always kept)
  bool force = false;          // -f (outputs are always overwritten)
  bool quiet = false;          // -q
  int verbose = 0;             // -v (repeatable)
  bool small = false;          // -s (accepted; no effect)
  std::vector<std::string> input_files;
};

// Parses argv like bzip2: short flags combine (-dc, -9v), long forms
// (--decompress, --stdout, ...), "--" ends the options.  A program name
// containing "bunzip" decompresses by default, one containing "bzcat"
// decompresses to standard output.  -h, -L and -V print and exit.
CliOptions parse_args(int argc, char **argv);

#endif  // BZG_CLI_H

// cli.cpp - option parsing (reference: src/utils/argsparce.cpp:6-101), with
// the bzip2 flags the reference's help lists as not yet available
// (argsparce.cpp:11-21).
#include "cli.h"

#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bzg.h"

static void show_help() {
  printf("bzip2 compressor for AMD MI355X (HIP)\n"
         "   usage: bzip2gpu [flags and input files in any order]\n"
         "\n"
         "   -h --help           print this message\n"
         "   -d --decompress     force decompression\n"
         "   -z --compress       force compression\n"
         "   -k --keep           keep (don't delete) input files (always the case)\n"
         "   -f --force          overwrite existing output files (always the case)\n"
         "   -t --test           test compressed file integrity\n"
         "   -c --stdout         output to standard out\n"
         "   -q --quiet          suppress noncritical error messages\n"
         "   -v --verbose        be verbose (a 2nd -v gives more)\n"
         "   -L --license        display software version & license\n"
         "   -V --version        display software version & license\n"
         "   -s --small          accepted for compatibility; no effect\n"
         "   -1 .. -9            set block size to 100k .. 900k\n"
         "   --fast              alias for -1\n"
         "   --best              alias for -9\n"
         "\n"
         "   If invoked as `bunzip2gpu', default action is to decompress;\n"
         "   as `bzcatgpu', to decompress to stdout.\n"
         "\n"
         "   If no file names are given, bzip2gpu compresses or decompresses\n"
         "   from standard input to standard output.  You can combine\n"
         "   short flags, so `-z -4' means the same as -z4 or -zv, &c.\n"
         "   Input files are kept and existing outputs are overwritten.\n");
}

static void show_version() {
  printf("%s\n"
         "   Reads and writes the bzip2 1.0.x stream format; compression cuts the\n"
         "   input into 900000-byte blocks as SebastianBoteroLeonik/bzip2gpu does.\n",
         bzg_version());
}

CliOptions parse_args(int argc, char **argv) {
  CliOptions options;
  if (argc > 0 && argv[0]) {
    const char *base = strrchr(argv[0], '/');
    base = base ? base + 1 : argv[0];
    if (strstr(base, "bunzip")) options.compress = false;
    if (strstr(base, "bzcat")) {
      options.compress = false;
      options.stdout_output = true;
    }
  }
  optind = 1;
  static struct option long_options[] = {
      {"help", no_argument, 0, 'h'},   {"decompress", no_argument, 0, 'd'}, {"compress", no_argument, 0, 'z'},
      {"keep", no_argument, 0, 'k'},   {"force", no_argument, 0, 'f'},      {"test", no_argument, 0, 't'},
      {"stdout", no_argument, 0, 'c'}, {"quiet", no_argument, 0, 'q'},      {"verbose", no_argument, 0, 'v'},
      {"license", no_argument, 0, 'L'}, {"version", no_argument, 0, 'V'},   {"small", no_argument, 0, 's'},
      {"fast", no_argument, 0, '1'},   {"best", no_argument, 0, '9'},       {0, 0, 0, 0}};
  for (;;) {
    int c = getopt_long(argc, argv, ":hdzkftcqvLVs123456789", long_options, NULL);
    if (c == -1) break;
    switch (c) {
      case 'z': options.compress = true; break;
      case 'd': options.compress = false; break;
      case 'k': options.keep = true; break;
      case 'f': options.force = true; break;
      case 't': options.test = true; options.compress = false; break;
      case 'c': options.stdout_output = true; break;
      case 'q': options.quiet = true; break;
      case 'v': options.verbose++; break;
      case 's': options.small = true; break;
      case '1': case '2': case '3': case '4': case '5': case '6': case '7': case '8': case '9':
        options.block_size = c - '0';
        break;
      case 'h': show_help(); exit(EXIT_SUCCESS);
      case 'L': case 'V': show_version(); exit(EXIT_SUCCESS);
      case ':': fprintf(stderr, "Missing argument\n"); exit(EXIT_FAILURE);
      case '?':
        printf("\n");
        if (optopt) printf("UNKNOWN OPTION: -%c\n", optopt);
        else printf("UNKNOWN OPTION: %s\n", argv[optind - 1]);
        printf("\n");
        show_help();
        exit(EXIT_FAILURE);
      default:
        fprintf(stderr, "?? getopt returned character code 0%o ??\n", c);
        exit(EXIT_FAILURE);
    }
  }
  while (optind < argc) options.input_files.push_back(argv[optind++]);
  return options;
}

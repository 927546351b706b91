// cli.cpp - option parsing (reference: src/utils/argsparce.cpp:6-101).
#include "cli.h"

#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>

static void show_help() {
  printf("bzip2 compressor for AMD MI355X (HIP)\n"
         "   usage: bzip2gpu [flags and input files in any order]\n"
         "\n"
         "   -h --help           print this message\n"
         "   -z --compress       force compression\n"
         "   -c --stdout         output to standard out\n"
         "   -1 .. -9            set block size to 100k .. 900k\n"
         "   --fast              alias for -1\n"
         "   --best              alias for -9\n"
         "\n"
         "   If no file names are given, bzip2gpu compresses from standard\n"
         "   input to standard output.  You can combine short flags, so\n"
         "   `-z -4' means the same as -z4.\n");
}

CliOptions parse_args(int argc, char **argv) {
  CliOptions options;
  optind = 1;
  static struct option long_options[] = {{"help", no_argument, 0, 'h'},   {"compress", no_argument, 0, 'z'},
                                         {"stdout", no_argument, 0, 'c'}, {"fast", no_argument, 0, '1'},
                                         {"best", no_argument, 0, '9'},   {0, 0, 0, 0}};
  for (;;) {
    int c = getopt_long(argc, argv, ":hzc123456789", long_options, NULL);
    if (c == -1) break;
    switch (c) {
      case 'z': options.compress = true; break;
      case 'c': options.stdout_output = true; break;
      case '1': case '2': case '3': case '4': case '5': case '6': case '7': case '8': case '9':
        options.block_size = c - '0';
        break;
      case 'h': show_help(); exit(EXIT_SUCCESS);
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

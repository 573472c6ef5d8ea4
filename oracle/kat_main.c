/* kat: prints the libmceliece test-driver KAT checksum of the oracle for one parameter set.
   usage: kat <set 0..4> <loops 8|64>   (test infrastructure only) */
#include <stdio.h>
#include <stdlib.h>
#include "tmce_oracle.h"
int main(int argc, char **argv) {
  if (argc < 3) { fprintf(stderr, "usage: kat set loops\n"); return 2; }
  char hex[65];
  oracle_kat_checksum(atoi(argv[1]), atoi(argv[2]), hex);
  printf("%s\n", hex);
  return 0;
}

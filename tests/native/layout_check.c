/* Host build against include/emo_api.h: the sizes and member offsets the
 * ctypes mirrors in apc_524_project_amd/_lib.py must reproduce. */
#include <stddef.h>
#include <stdio.h>

#include "../../include/emo_api.h"

int main(void) {
  printf("emo_config %zu max_particles %zu neutralize %zu waist %zu\n", sizeof(emo_config),
         offsetof(emo_config, max_particles), offsetof(emo_config, neutralize), offsetof(emo_config, waist));
  printf("emo_loader %zu drift %zu seed %zu\n", sizeof(emo_loader), offsetof(emo_loader, drift),
         offsetof(emo_loader, seed));
  return 0;
}

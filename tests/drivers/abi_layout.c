/* prints the sizes and offsets the reference's callers depend on, as compiled against include/mlib.h */
#include <stddef.h>
#include <stdio.h>
#include "mlib.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(ANN), sizeof(PRS), sizeof(LNK), sizeof(Cell), offsetof(LNK, B),
         offsetof(Cell, p), _Alignof(Cell));
  return 0;
}

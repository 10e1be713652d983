// C-callable entry to the reference's ludcmp (src/utils/ludcmp.C:9), for checking the
// oracle's LU bit for bit.  Built only into oracle/_ref, never shipped with the product.
#include <complex>
using namespace std;
#include "config.h"
#include "ludcmp.h"

extern "C" void ref_ludcmp(double* a, int n, int* indx, double* d) {
  ludcmp(reinterpret_cast<complex<Float>*>(a), n, indx, d);
}

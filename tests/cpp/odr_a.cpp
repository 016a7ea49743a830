// ODR check: the header-only templates included in two translation units
// linked into one program (the reference's test/extra/test_multiple_objects*).
#include <boost/compute.hpp>
int odr_b();
int main() { return odr_b() == 7 ? 0 : 1; }

#include <boost/compute.hpp>
int odr_b()
{
    boost::compute::plus<int> p;
    return p(3, 4);
}

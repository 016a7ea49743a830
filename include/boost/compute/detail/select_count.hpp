//---------------------------------------------------------------------------//
// The number of elements a selection call wrote, counted on the device and
// read back (blocking), as transform_if_impl reads the last scanned index
// (reference: algorithm/transform_if.hpp:57-62).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_DETAIL_SELECT_COUNT_HPP
#define BOOST_COMPUTE_DETAIL_SELECT_COUNT_HPP

#include <cstddef>
#include <cstdint>

#include <boost/compute/container/vector.hpp>

namespace boost {
namespace compute {
namespace detail {

class select_count {
public:
    explicit select_count(command_queue &queue) : m_v(1, queue.get_context()) {}
    uint64_t *ptr() { return (uint64_t *)m_v.begin().device_ptr(); }
    std::ptrdiff_t read(command_queue &queue) { return (std::ptrdiff_t)m_v.begin().read(queue); }

private:
    ::boost::compute::vector<unsigned long long> m_v;
};

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif

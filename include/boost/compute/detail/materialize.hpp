//---------------------------------------------------------------------------//
// A generated or mapped input range written out to a temporary vector (one
// bcmi_transform_iter launch), for the algorithms whose kernels read plain
// arrays only (reduce / accumulate over counting, constant and
// transform iterators; the reference's general inner_product likewise
// transforms into a temporary first, inner_product.hpp).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_DETAIL_MATERIALIZE_HPP
#define BOOST_COMPUTE_DETAIL_MATERIALIZE_HPP

#include <boost/compute/container/vector.hpp>
#include <boost/compute/detail/iter_desc.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class InputIterator>
inline ::boost::compute::vector<typename iter_desc<InputIterator>::value_type> materialize(
    InputIterator first, size_t n, command_queue &queue)
{
    typedef typename iter_desc<InputIterator>::value_type T;
    ::boost::compute::vector<T> tmp(n, queue.get_context());
    const bcmi_iter d = iter_desc<InputIterator>::get(first);
    if (n)
        bcmi_check(bcmi_transform_iter(queue.get(), bcmi_dtype<T>(), 1, BCMI_PLUS, &d, 0, tmp.begin().device_ptr(),
                                       n));
    return tmp;
}

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif

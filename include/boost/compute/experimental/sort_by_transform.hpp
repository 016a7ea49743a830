//---------------------------------------------------------------------------//
// experimental::sort_by_transform (reference:
// experimental/sort_by_transform.hpp): keys = transform(element) into a
// temporary, then sort_by_key(keys, elements, compare) -- the same steps,
// on a key-extraction kernel and the radix sort_by_key with the elements as
// its values (1, 2, 4, 8 or 16 bytes).  transform: abs<T> on scalar
// elements, or get<N> on vector-type elements; compare: less or greater.
// Other transforms (length, user functions) do not compile.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_EXPERIMENTAL_SORT_BY_TRANSFORM_HPP
#define BOOST_COMPUTE_EXPERIMENTAL_SORT_BY_TRANSFORM_HPP

#include <cstddef>
#include <iterator>
#include <type_traits>

#include <bcmi.h>

#include <boost/compute/container/vector.hpp>
#include <boost/compute/detail/bcmi.hpp>
#include <boost/compute/functional/get.hpp>
#include <boost/compute/functional/integer.hpp>
#include <boost/compute/functional/operator.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/types/fundamental.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T, class Transform>
struct key_extraction {
    static_assert(sizeof(Transform) == 0, "sort_by_transform: transform not recognised (abs<T>, get<N>)");
};

template <class T, class U>
struct key_extraction<T, ::boost::compute::abs<U> > {
    typedef T key_type;
    static const int fn = BCMI_MAP_ABS;
    static std::size_t offset() { return 0; }
};

template <class T, std::size_t K>
struct key_extraction<T, ::boost::compute::get<K> > {
    typedef typename T::scalar_type key_type;
    static_assert(K < T::size, "get<N>: no such component");
    static const int fn = BCMI_MAP_ID;
    static std::size_t offset() { return K * sizeof(key_type); }
};

template <class Compare>
struct by_transform_direction {
    static_assert(sizeof(Compare) == 0, "sort_by_transform: compare must be less or greater");
};
template <class X>
struct by_transform_direction<less<X> > {
    static const int descending = 0;
};
template <class X>
struct by_transform_direction<greater<X> > {
    static const int descending = 1;
};

}  // namespace detail

namespace experimental {

template <class Iterator, class Transform, class Compare>
inline void sort_by_transform(Iterator first, Iterator last, Transform, Compare,
                              command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<Iterator>::value_type T;
    typedef detail::key_extraction<T, Transform> X;
    typedef typename X::key_type K;
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16,
                  "sort_by_transform: elements of 1, 2, 4, 8 or 16 bytes");
    const std::size_t n = (std::size_t)(last - first);
    if (n < 2) return;
    ::boost::compute::vector<K> keys(n, queue.get_context());
    detail::bcmi_check(bcmi_extract_keys(queue.get(), detail::bcmi_dtype<K>(), X::fn, first.device_ptr(), sizeof(T),
                                         X::offset(), n, keys.begin().device_ptr()));
    detail::bcmi_check(bcmi_sort(queue.get(), BCMI_SORT_BY_KEY, detail::bcmi_dtype<K>(), keys.begin().device_ptr(),
                                 first.device_ptr(), sizeof(T), n, detail::by_transform_direction<Compare>::descending));
}

}  // namespace experimental
}  // namespace compute
}  // namespace boost

#endif

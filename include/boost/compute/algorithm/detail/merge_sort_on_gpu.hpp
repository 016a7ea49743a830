//---------------------------------------------------------------------------//
// detail::merge_sort_on_gpu / merge_sort_by_key_on_gpu (reference:
// algorithm/detail/merge_sort_on_gpu.hpp:467-598): the comparison sort that
// sort, sort_by_key, stable_sort and stable_sort_by_key use for a user
// comparator (sort.hpp:84-104, sort_by_key.hpp:88-106, stable_sort.hpp:36-50).
//
// A comparator is one of the lambda forms of lambda.hpp (_1 < _2,
// get<N>(_1) < get<M>(_2), abs, abs of a sum, length; < or >) or a struct
// member comparison, experimental::less_by(&S::m) / greater_by(&S::m) (the
// reference's BOOST_COMPUTE_FUNCTION(bool, f, (S a, S b), { return a.m <
// b.m; })).  Each becomes a bcmi_comparator, and bcmi_merge_sort sorts the
// elements, stable at every n (the reference's unstable merge sort may
// order ties differently; its n <= 32 insertion sort is stable too).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_DETAIL_MERGE_SORT_ON_GPU_HPP
#define BOOST_COMPUTE_ALGORITHM_DETAIL_MERGE_SORT_ON_GPU_HPP

#include <cstddef>
#include <type_traits>

#include <bcmi.h>

#include <boost/compute/algorithm/detail/radix_sort.hpp>
#include <boost/compute/detail/bcmi.hpp>
#include <boost/compute/lambda.hpp>
#include <boost/compute/system.hpp>

namespace boost {
namespace compute {
namespace experimental {

// a.m < b.m (or >) for a member m of a fundamental type
template <class R, class F>
struct member_compare {
    std::size_t offset;
    bool greater;
};

template <class R, class F>
inline std::size_t member_offset(F R::*m)
{
    static_assert(std::is_standard_layout<R>::value, "less_by/greater_by: the record must be standard-layout");
    typename std::aligned_storage<sizeof(R), alignof(R)>::type buf;
    const R *r = reinterpret_cast<const R *>(&buf);
    return (std::size_t)(reinterpret_cast<const char *>(&(r->*m)) - reinterpret_cast<const char *>(r));
}

template <class R, class F>
inline member_compare<R, F> less_by(F R::*m)
{
    return member_compare<R, F>{member_offset(m), false};
}

template <class R, class F>
inline member_compare<R, F> greater_by(F R::*m)
{
    return member_compare<R, F>{member_offset(m), true};
}

}  // namespace experimental

namespace detail {

template <class C>
struct is_comparator_expr : std::false_type {
};
template <int F, std::size_t L, std::size_t R, std::size_t S, bool G>
struct is_comparator_expr<lambda::compare_expr<F, L, R, S, G> > : std::true_type {
};
template <class R, class F>
struct is_comparator_expr<experimental::member_compare<R, F> > : std::true_type {
};

// the scalar type and component count of an element (a vector type or a scalar)
template <class T, class = void>
struct record_layout {
    typedef T scalar_type;
    static const std::size_t size = 1;
};
template <class T>
struct record_layout<T, typename std::conditional<false, typename T::scalar_type, void>::type> {
    typedef typename T::scalar_type scalar_type;
    static const std::size_t size = T::size;
};

template <class T, int F, std::size_t L, std::size_t R, std::size_t S, bool G>
inline bcmi_comparator make_comparator(lambda::compare_expr<F, L, R, S, G>)
{
    typedef record_layout<T> lay;
    typedef typename lay::scalar_type scalar;
    static_assert(L < lay::size && R < lay::size && S < lay::size, "comparator: no such component");
    static_assert(F != BCMI_KEY_LENGTH || std::is_floating_point<scalar>::value,
                  "length() compares float or double vectors");
    bcmi_comparator c;
    c.dtype = bcmi_dtype<scalar>();
    c.form = F;
    c.lhs_offset = (uint32_t)(L * sizeof(scalar));
    c.rhs_offset = (uint32_t)(R * sizeof(scalar));
    c.offset2 = (uint32_t)(S * sizeof(scalar));
    c.count = F == BCMI_KEY_LENGTH ? (uint32_t)lay::size : 0u;
    c.greater = G ? 1 : 0;
    return c;
}

template <class T, class R, class F>
inline bcmi_comparator make_comparator(experimental::member_compare<R, F> m)
{
    static_assert(std::is_same<T, R>::value, "less_by/greater_by: member of another type");
    bcmi_comparator c;
    c.dtype = bcmi_dtype<F>();
    c.form = BCMI_KEY_FIELD;
    c.lhs_offset = c.rhs_offset = (uint32_t)m.offset;
    c.offset2 = 0;
    c.count = 0;
    c.greater = m.greater ? 1 : 0;
    return c;
}

// less<T> / greater<T> passed to merge_sort_on_gpu directly (as the
// reference's own tests do, test/test_merge_sort_gpu.cpp:38-62)
template <class T>
inline bcmi_comparator scalar_comparator(bool greater)
{
    static_assert(std::is_arithmetic<T>::value, "less/greater comparator of a non-scalar type");
    bcmi_comparator c;
    c.dtype = bcmi_dtype<T>();
    c.form = BCMI_KEY_FIELD;
    c.lhs_offset = c.rhs_offset = c.offset2 = c.count = 0;
    c.greater = greater ? 1 : 0;
    return c;
}

template <class T, class X>
inline bcmi_comparator make_comparator(less<X>)
{
    return scalar_comparator<T>(false);
}

template <class T, class X>
inline bcmi_comparator make_comparator(greater<X>)
{
    return scalar_comparator<T>(true);
}

template <class T, class V, class Compare>
inline void comparator_sort_call(buffer_iterator<T> first, buffer_iterator<T> last, V *values, Compare compare,
                                 command_queue &queue)
{
    const bcmi_comparator c = make_comparator<T>(compare);
    bcmi_check(bcmi_merge_sort(queue.get(), &c, first.device_ptr(), sizeof(T), values, values ? sizeof(V) : 0,
                               (size_t)(last - first)));
}

// The device sort entry of the sort templates: comparator expressions go to
// the comparison sort, less<T> / greater<T> to the radix dispatch.
template <class K, class V, class Compare>
inline void sort_call(int api, buffer_iterator<K> first, buffer_iterator<K> last, V *values, Compare compare,
                      command_queue &queue, std::true_type)
{
    comparator_sort_call(first, last, values, compare, queue);
}

template <class K, class V, class Compare>
inline void sort_call(int api, buffer_iterator<K> first, buffer_iterator<K> last, V *values, Compare,
                      command_queue &queue, std::false_type)
{
    bcmi_sort_call<K, V>(api, first, last, values, sort_direction<Compare, K>::descending, queue);
}

template <class K, class V, class Compare>
inline void sort_call(int api, buffer_iterator<K> first, buffer_iterator<K> last, V *values, Compare compare,
                      command_queue &queue)
{
    sort_call(api, first, last, values, compare, queue,
              std::integral_constant<bool, is_comparator_expr<Compare>::value>());
}

template <class T, class Compare>
inline void merge_sort_on_gpu(buffer_iterator<T> first, buffer_iterator<T> last, Compare compare, bool /*stable*/,
                              command_queue &queue)
{
    comparator_sort_call(first, last, (T *)0, compare, queue);
}

template <class T, class Compare>
inline void merge_sort_on_gpu(buffer_iterator<T> first, buffer_iterator<T> last, Compare compare,
                              command_queue &queue)
{
    comparator_sort_call(first, last, (T *)0, compare, queue);
}

template <class K, class V, class Compare>
inline void merge_sort_by_key_on_gpu(buffer_iterator<K> keys_first, buffer_iterator<K> keys_last,
                                     buffer_iterator<V> values_first, Compare compare, bool /*stable*/,
                                     command_queue &queue)
{
    comparator_sort_call(keys_first, keys_last, values_first.device_ptr(), compare, queue);
}

template <class K, class V, class Compare>
inline void merge_sort_by_key_on_gpu(buffer_iterator<K> keys_first, buffer_iterator<K> keys_last,
                                     buffer_iterator<V> values_first, Compare compare, command_queue &queue)
{
    comparator_sort_call(keys_first, keys_last, values_first.device_ptr(), compare, queue);
}

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif

//---------------------------------------------------------------------------//
// transform_iterator (reference: iterator/transform_iterator.hpp:33-227):
// a base iterator seen through a unary function.  The reference prints the
// function into the consuming kernel; here the function must be one of the
// element maps the kernels implement (bcmi_iter's fn):
//   _1 - c, _1 + c, abs<T>(), abs(_1), _1 * _1, pow(_1, 2), pown(_1, 2),
//   pow(_1 - c, 2), and the comparisons _1 < c, _1 > c, ... (1 or 0),
// over a buffer_iterator, counting_iterator or constant_iterator base.
// Anything else does not compile (detail::element_map_of).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ITERATOR_TRANSFORM_ITERATOR_HPP
#define BOOST_COMPUTE_ITERATOR_TRANSFORM_ITERATOR_HPP

#include <cstddef>
#include <iterator>
#include <type_traits>

#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class UnaryFunction>
class transform_iterator {
public:
    typedef typename std::iterator_traits<InputIterator>::value_type base_value_type;
    // the reference's value_type is the function's result type; every map
    // here keeps the element type (a comparison's int result converts back)
    typedef base_value_type value_type;
    typedef std::ptrdiff_t difference_type;
    typedef const value_type *pointer;
    typedef value_type reference;
    typedef std::random_access_iterator_tag iterator_category;

    transform_iterator(InputIterator it, UnaryFunction f) : m_it(it), m_f(f) {}

    const InputIterator &base() const { return m_it; }
    const UnaryFunction &functor() const { return m_f; }

    transform_iterator operator+(std::ptrdiff_t d) const { return transform_iterator(m_it + d, m_f); }
    transform_iterator operator-(std::ptrdiff_t d) const { return transform_iterator(m_it - d, m_f); }
    std::ptrdiff_t operator-(const transform_iterator &o) const { return m_it - o.m_it; }
    transform_iterator &operator++()
    {
        m_it = m_it + 1;
        return *this;
    }
    bool operator==(const transform_iterator &o) const { return m_it == o.m_it; }
    bool operator!=(const transform_iterator &o) const { return !(*this == o); }

private:
    InputIterator m_it;
    UnaryFunction m_f;
};

template <class InputIterator, class UnaryFunction>
inline transform_iterator<InputIterator, UnaryFunction> make_transform_iterator(InputIterator it, UnaryFunction f)
{
    return transform_iterator<InputIterator, UnaryFunction>(it, f);
}

template <class InputIterator, class UnaryFunction>
struct is_device_iterator<transform_iterator<InputIterator, UnaryFunction> > : std::true_type {};

}  // namespace compute
}  // namespace boost

#endif

//---------------------------------------------------------------------------//
// constant_iterator (reference: iterator/constant_iterator.hpp): the value
// v repeated; nothing is stored.  The algorithms pass it to the device as a
// generated sequence (bcmi_iter, BCMI_ITER_CONSTANT).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ITERATOR_CONSTANT_ITERATOR_HPP
#define BOOST_COMPUTE_ITERATOR_CONSTANT_ITERATOR_HPP

#include <cstddef>
#include <iterator>
#include <type_traits>

#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class T>
class constant_iterator {
public:
    typedef T value_type;
    typedef std::ptrdiff_t difference_type;
    typedef const T *pointer;
    typedef T reference;
    typedef std::random_access_iterator_tag iterator_category;

    explicit constant_iterator(const T &value = T(), std::size_t index = 0) : m_value(value), m_index(index) {}

    T value() const { return m_value; }
    std::size_t index() const { return m_index; }
    T operator*() const { return m_value; }

    constant_iterator operator+(std::ptrdiff_t d) const { return constant_iterator(m_value, m_index + d); }
    constant_iterator operator-(std::ptrdiff_t d) const { return constant_iterator(m_value, m_index - d); }
    std::ptrdiff_t operator-(const constant_iterator &o) const
    {
        return (std::ptrdiff_t)m_index - (std::ptrdiff_t)o.m_index;
    }
    constant_iterator &operator++() { ++m_index; return *this; }
    bool operator==(const constant_iterator &o) const { return m_index == o.m_index; }
    bool operator!=(const constant_iterator &o) const { return !(*this == o); }

private:
    T m_value;
    std::size_t m_index;
};

template <class T>
inline constant_iterator<T> make_constant_iterator(const T &value, std::size_t index = 0)
{
    return constant_iterator<T>(value, index);
}

template <class T>
struct is_device_iterator<constant_iterator<T> > : std::true_type {};

}  // namespace compute
}  // namespace boost

#endif

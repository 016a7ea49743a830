//---------------------------------------------------------------------------//
// counting_iterator (reference: iterator/counting_iterator.hpp): the
// sequence init, init + 1, ...; nothing is stored.  The reductions here
// (transform_reduce, inner_product) generate it on the device.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ITERATOR_COUNTING_ITERATOR_HPP
#define BOOST_COMPUTE_ITERATOR_COUNTING_ITERATOR_HPP

#include <cstddef>
#include <iterator>
#include <type_traits>

#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class T>
class counting_iterator {
public:
    typedef T value_type;
    typedef std::ptrdiff_t difference_type;
    typedef const T *pointer;
    typedef T reference;
    typedef std::random_access_iterator_tag iterator_category;

    explicit counting_iterator(const T &init = T()) : m_init(init), m_index(0) {}
    counting_iterator(const T &init, std::ptrdiff_t index) : m_init(init), m_index(index) {}

    T init() const { return m_init; }
    std::ptrdiff_t index() const { return m_index; }
    T operator*() const { return (T)(m_init + (T)m_index); }

    counting_iterator operator+(std::ptrdiff_t d) const { return counting_iterator(m_init, m_index + d); }
    counting_iterator operator-(std::ptrdiff_t d) const { return counting_iterator(m_init, m_index - d); }
    std::ptrdiff_t operator-(const counting_iterator &o) const
    {
        return (std::ptrdiff_t)(m_init + (T)m_index) - (std::ptrdiff_t)(o.m_init + (T)o.m_index);
    }
    counting_iterator &operator++() { ++m_index; return *this; }
    bool operator==(const counting_iterator &o) const { return *(*this) == *o; }
    bool operator!=(const counting_iterator &o) const { return !(*this == o); }

private:
    T m_init;
    std::ptrdiff_t m_index;
};

template <class T>
inline counting_iterator<T> make_counting_iterator(const T &init)
{
    return counting_iterator<T>(init);
}

template <class T>
struct is_device_iterator<counting_iterator<T> > : std::true_type {};

}  // namespace compute
}  // namespace boost

#endif

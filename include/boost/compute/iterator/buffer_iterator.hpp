//---------------------------------------------------------------------------//
// buffer_iterator<T>: random-access iterator over a buffer, holding the
// buffer and an element index (reference: iterator/buffer_iterator.hpp:
// 136-190).  Algorithms turn (buffer, index) into one element pointer, so no
// offset is ever compiled into a kernel.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ITERATOR_BUFFER_ITERATOR_HPP
#define BOOST_COMPUTE_ITERATOR_BUFFER_ITERATOR_HPP

#include <cstddef>
#include <iterator>

#include <boost/compute/buffer.hpp>
#include <boost/compute/command_queue.hpp>

namespace boost {
namespace compute {

template <class T>
class buffer_iterator
{
public:
    typedef T value_type;
    typedef std::ptrdiff_t difference_type;
    typedef T *pointer;
    typedef T &reference;
    typedef std::random_access_iterator_tag iterator_category;

    buffer_iterator() : m_index(0) {}
    buffer_iterator(const buffer &b, size_t index) : m_buffer(b), m_index(index) {}

    const buffer &get_buffer() const { return m_buffer; }
    size_t get_index() const { return m_index; }
    // raw device address of this element
    T *device_ptr() const { return static_cast<T *>(m_buffer.get()) + m_index; }

    // blocking single-element access (buffer_iterator::read/write)
    T read(command_queue &q) const
    {
        T v;
        detail::bcmi_check(bcmi_memcpy(q.get(), &v, device_ptr(), sizeof(T), BCMI_D2H, 1));
        return v;
    }
    void write(const T &v, command_queue &q) const
    {
        detail::bcmi_check(bcmi_memcpy(q.get(), device_ptr(), &v, sizeof(T), BCMI_H2D, 1));
    }

    buffer_iterator &operator++() { ++m_index; return *this; }
    buffer_iterator operator++(int) { buffer_iterator t(*this); ++m_index; return t; }
    buffer_iterator &operator--() { --m_index; return *this; }
    buffer_iterator operator--(int) { buffer_iterator t(*this); --m_index; return t; }
    buffer_iterator &operator+=(difference_type n) { m_index += n; return *this; }
    buffer_iterator &operator-=(difference_type n) { m_index -= n; return *this; }
    buffer_iterator operator+(difference_type n) const { return buffer_iterator(m_buffer, m_index + n); }
    buffer_iterator operator-(difference_type n) const { return buffer_iterator(m_buffer, m_index - n); }
    difference_type operator-(const buffer_iterator &o) const
    {
        return (difference_type)m_index - (difference_type)o.m_index;
    }
    bool operator==(const buffer_iterator &o) const { return m_buffer == o.m_buffer && m_index == o.m_index; }
    bool operator!=(const buffer_iterator &o) const { return !(*this == o); }
    bool operator<(const buffer_iterator &o) const { return m_index < o.m_index; }
    bool operator>(const buffer_iterator &o) const { return m_index > o.m_index; }
    bool operator<=(const buffer_iterator &o) const { return m_index <= o.m_index; }
    bool operator>=(const buffer_iterator &o) const { return m_index >= o.m_index; }

private:
    buffer m_buffer;
    size_t m_index;
};

template <class T>
inline buffer_iterator<T> make_buffer_iterator(const buffer &b, size_t index = 0)
{
    return buffer_iterator<T>(b, index);
}

}  // namespace compute
}  // namespace boost

#endif

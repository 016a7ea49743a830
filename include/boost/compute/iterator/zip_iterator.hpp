//---------------------------------------------------------------------------//
// zip_iterator (reference: iterator/zip_iterator.hpp:40-316): several
// ranges advanced together, element i being the tuple of their elements.
// The reference takes a boost::tuple; this build has no Boost and takes a
// std::tuple (make_zip_iterator(std::make_tuple(a, b))).  The algorithms
// read one component at a time: transform(zip, zip_end, out, get<N>())
// copies component N's range (test/test_zip_iterator.cpp:157-221).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ITERATOR_ZIP_ITERATOR_HPP
#define BOOST_COMPUTE_ITERATOR_ZIP_ITERATOR_HPP

#include <cstddef>
#include <iterator>
#include <tuple>
#include <type_traits>

#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class IteratorTuple>
class zip_iterator {
public:
    typedef std::ptrdiff_t difference_type;
    typedef std::random_access_iterator_tag iterator_category;

    explicit zip_iterator(const IteratorTuple &its) : m_its(its) {}

    const IteratorTuple &get_iterator_tuple() const { return m_its; }

    zip_iterator operator+(std::ptrdiff_t d) const { return zip_iterator(advance(m_its, d)); }
    std::ptrdiff_t operator-(const zip_iterator &o) const { return std::get<0>(m_its) - std::get<0>(o.m_its); }
    bool operator==(const zip_iterator &o) const { return std::get<0>(m_its) == std::get<0>(o.m_its); }
    bool operator!=(const zip_iterator &o) const { return !(*this == o); }

private:
    template <std::size_t... I>
    static IteratorTuple advance_(const IteratorTuple &t, std::ptrdiff_t d, std::index_sequence<I...>)
    {
        return IteratorTuple((std::get<I>(t) + d)...);
    }
    static IteratorTuple advance(const IteratorTuple &t, std::ptrdiff_t d)
    {
        return advance_(t, d, std::make_index_sequence<std::tuple_size<IteratorTuple>::value>());
    }

    IteratorTuple m_its;
};

template <class IteratorTuple>
inline zip_iterator<IteratorTuple> make_zip_iterator(const IteratorTuple &its)
{
    return zip_iterator<IteratorTuple>(its);
}

template <class IteratorTuple>
struct is_device_iterator<zip_iterator<IteratorTuple> > : std::true_type {};

}  // namespace compute
}  // namespace boost

#endif

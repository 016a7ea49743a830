//---------------------------------------------------------------------------//
// not1 / unary_negate (reference: functional/logical.hpp:113-175): negates
// a recognised unary predicate (remove_if and stable_partition's second
// copy_if use it).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_FUNCTIONAL_LOGICAL_HPP
#define BOOST_COMPUTE_FUNCTIONAL_LOGICAL_HPP

namespace boost {
namespace compute {

template <class Predicate>
class unary_negate {
public:
    typedef int result_type;
    explicit unary_negate(Predicate pred) : m_pred(pred) {}
    const Predicate &predicate() const { return m_pred; }

private:
    Predicate m_pred;
};

template <class Predicate>
inline unary_negate<Predicate> not1(const Predicate &predicate)
{
    return unary_negate<Predicate>(predicate);
}

}  // namespace compute
}  // namespace boost

#endif

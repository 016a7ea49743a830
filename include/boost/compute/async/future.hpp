//---------------------------------------------------------------------------//
// future<T> (reference: async/future.hpp): a result that is valid once an
// event has completed.  get() waits, then returns the result.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ASYNC_FUTURE_HPP
#define BOOST_COMPUTE_ASYNC_FUTURE_HPP

#include <boost/compute/event.hpp>

namespace boost {
namespace compute {

template <class T>
class future {
public:
    future() {}
    future(const T &result, const event &ev) : m_result(result), m_event(ev) {}
    T get()
    {
        wait();
        return m_result;
    }
    bool valid() const { return m_event.valid(); }
    void wait() const { m_event.wait(); }
    event get_event() const { return m_event; }

private:
    T m_result;
    event m_event;
};

}  // namespace compute
}  // namespace boost

#endif

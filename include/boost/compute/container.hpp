#ifndef BOOST_COMPUTE_CONTAINER_HPP
#define BOOST_COMPUTE_CONTAINER_HPP
#include <boost/compute/container/mapped_view.hpp>
#include <boost/compute/container/vector.hpp>
#endif

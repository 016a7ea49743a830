//---------------------------------------------------------------------------//
// Umbrella header of the MI355X-native boost::compute hot path: the same
// include as the reference's <boost/compute.hpp> for the sort / scan /
// reduce / reduce_by_key / transform primitives.  Link with -lbcmi.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_HPP
#define BOOST_COMPUTE_HPP
#include <boost/compute/algorithm.hpp>
#include <boost/compute/algorithm/detail/merge_sort_on_gpu.hpp>
#include <boost/compute/algorithm/detail/radix_sort.hpp>
#include <boost/compute/async/future.hpp>
#include <boost/compute/container.hpp>
#include <boost/compute/core.hpp>
#include <boost/compute/experimental/sort_by_transform.hpp>
#include <boost/compute/functional.hpp>
#include <boost/compute/iterator.hpp>
#include <boost/compute/lambda.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>
#endif

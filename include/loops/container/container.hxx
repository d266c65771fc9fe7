/**
 * @file container.hxx
 * @brief Every container header at once (reference include/loops/container/
 * container.hxx:12-14): the sparse formats, the Matrix Market reader, the
 * vectors and the dense matrix.
 */
#pragma once

#include <loops/container/formats.hxx>
#include <loops/container/market.hxx>
#include <loops/container/matrix.cuh>
#include <loops/container/vector.hxx>

// array_data.hpp -- ArrayData (include/data/array_data.h:9-29): one named array handed
// from a particle to the I/O layer, 1-D or 2-D (x, y columns), double or long.  Here the
// vectors are std::vector and per-particle data is already in the static particle order
// (the order of the first upload), which is the order the reference's StateLog writes
// after its scatter by static_particle_index (io_utils.cu:174-231).
#pragma once
#include <array>
#include <string>
#include <utility>
#include <variant>
#include <vector>

enum class DataType { Double, Long };

struct ArrayData {
    std::string name = "NULL";
    DataType type = DataType::Double;
    std::array<long, 2> size = {0, 0};  // {rows, 1 or 2 columns}
    std::string index_array_name;       // the reference's reorder index for this array
    std::variant<std::vector<double>, std::vector<long>,
                 std::pair<std::vector<double>, std::vector<double>>,
                 std::pair<std::vector<long>, std::vector<long>>>
        data;
};

// uppaal.hpp — UPPAAL XML front end.
#pragma once

#include <memory>
#include <string>
#include "network.hpp"

namespace smc {

struct Parsed {
    Network net;
    smc_network view_cache{};
    int goal_count = 0;        // goal locations in the network
    int query_goal_count = 0;  // of which named "Goal" or matched by the query
};

// Throw std::runtime_error with a message on any failure.
std::unique_ptr<Parsed> parse_uppaal_string(const std::string &xml, const std::string &loc_query);
std::unique_ptr<Parsed> parse_uppaal_file(const std::string &path, const std::string &loc_query);

}  // namespace smc

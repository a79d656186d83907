// `cuts <data_graph> <query_graph> [trunks]` -- the reference's command line
// (main.cu:24-46).  Two positional arguments run search(); a third runs
// search_dfs_bfs_strategy() with that many level-0 trunks (0 = automatic,
// where the reference would divide by zero).  Optional flags, accepted
// anywhere after the program name:
//   --write   also write ans.txt (the reference hard-codes write_to_disk=false)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/cuts.h"

int main(int argc, char* argv[]) {
  std::vector<std::string> pos;
  bool write = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--write")) write = true;
    else pos.push_back(argv[i]);
  }
  if (pos.size() < 2) {
    std::cout << "args data_graph path,query_graph_path,first depth trunks(optional)" << std::endl;
    exit(-1);
  }
  const std::string data = pos[0], query = pos[1];
  uint64_t count = 0;
  int rc;
  if (pos.size() == 2) {
    rc = cuts_search(query.c_str(), data.c_str(), write, &count);
  } else {
    const int t = atoi(pos[2].c_str());
    rc = cuts_search_dfs_bfs_strategy(query.c_str(), data.c_str(), write, t < 0 ? 0u : (uint32_t)t, &count);
  }
  if (rc != CUTS_OK) {
    std::cout << cuts_last_error() << std::endl;
    exit(-1);
  }
  return 0;
}

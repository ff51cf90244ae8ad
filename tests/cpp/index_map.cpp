// Index-map harness: prints, for a mesh configuration given on the command
// line, every cell's centre and the offset of its first dof, in the order
// of the library's vectors (the reference's DoFHandler numbering on the
// refine_global mesh). Runs on the host only; the same listing printed
// from a deal.II build of the reference (cell->center(), and
// cell->get_dof_indices()[0]) makes the numbering checkable line by line.
//
//   index_map dim degree refine [perturbation seed]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dgops.h"

int
main(int argc, char **argv)
{
  if (argc < 4)
    {
      std::fprintf(stderr, "usage: %s dim degree refine [perturbation seed]\n",
                   argv[0]);
      return 2;
    }
  dg_config cfg;
  dg_config_default(&cfg);
  cfg.dim          = std::atoi(argv[1]);
  cfg.degree       = std::atoi(argv[2]);
  cfg.refine       = std::atoi(argv[3]);
  cfg.perturbation = argc > 4 ? std::atof(argv[4]) : 0.;
  cfg.seed         = argc > 5 ? std::strtoull(argv[5], nullptr, 10) : 42;
  int64_t n = 0;
  if (dg_host_index_map(&cfg, &n, nullptr, nullptr) != DG_OK)
    {
      std::fprintf(stderr, "index_map: %s\n", dg_last_error());
      return 1;
    }
  std::vector<double>  centres(n * cfg.dim);
  std::vector<int64_t> offsets(n);
  if (dg_host_index_map(&cfg, &n, centres.data(), offsets.data()) != DG_OK)
    {
      std::fprintf(stderr, "index_map: %s\n", dg_last_error());
      return 1;
    }
  std::printf("# cell centre[0..dim) first_dof\n");
  for (int64_t c = 0; c < n; ++c)
    {
      std::printf("%lld", (long long)c);
      for (int d = 0; d < cfg.dim; ++d)
        std::printf(" %.17g", centres[c * cfg.dim + d]);
      std::printf(" %lld\n", (long long)offsets[c]);
    }
  return 0;
}

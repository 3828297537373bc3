"""python -m everest_amd.system.batch_query <graph> <gpu> <query-dir>
Runs every query of a directory (sorted) on one resident graph
(reference: system/batch-query.py:1-19)."""
import os
import sys

from everest_amd.tmine import frontend, locations


def main(graph, gpu, queryDirectory):
    queries = sorted(os.listdir(queryDirectory))
    m = frontend.Mining(os.path.join(locations.inputGraphRoot(), graph), int(gpu))
    for q in queries:
        print("@@@ [output starts]")
        print("# process %s" % q)
        print("@@@ [output ends]")
        m.run(os.path.join(queryDirectory, q))


if __name__ == "__main__":
    main(*sys.argv[1:4])

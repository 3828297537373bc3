"""python -m everest_amd.system.single_query <query.yaml>
(reference: system/single-query.py:1-10)."""
import sys

from everest_amd.tmine import frontend


def runQueryOnce(query):
    frontend.MiningSingle(query).run()


if __name__ == "__main__":
    runQueryOnce(sys.argv[1])

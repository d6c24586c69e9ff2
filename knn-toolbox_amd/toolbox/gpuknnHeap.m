function [dist, idx, time, dTime, sTime] = gpuknnHeap(query, pdata, k, numStreams)
% GPUKNNHEAP  k nearest neighbours on the GPU, heap entry point.
%   [dist idx time dTime sTime] = gpuknnHeap(query, pdata, k, numStreams)
%   query: d x q, one query per column; pdata: d x N, one corpus vector per column.
%   dist, idx: k x q squared Euclidean distances and 1-based column indices of pdata, each
%   column sorted by distance (ties: lower index first). time, dTime, sTime: seconds of the
%   call, its distance stage and its search stage. numStreams is accepted for compatibility.
%   Same calling convention as toolbox/gpuknnHeap.m of the CUDA toolbox.
if nargin < 4
  numStreams = 1;
end
if isa(pdata, 'single')
  [dist, idx, time, dTime, sTime] = mexknnsHeap(single(query), pdata, k, numStreams);
else
  [dist, idx, time, dTime, sTime] = mexknnsDHeap(double(query), double(pdata), k, numStreams);
end
[dist, idx] = knns_finish(dist, idx, k);
end

function [dist, idx] = knns_finish(dist, idx, k)
% 0-based indices -> 1-based; stable sort of every column by distance; first k rows
idx = idx + 1;
[dist, order] = sort(dist, 1);
for c = 1:size(idx, 2)
  idx(:, c) = idx(order(:, c), c);
end
dist = dist(1:k, :);
idx = idx(1:k, :);
end
